"""Robot builders: the oracle restatement against the reference's own dumped
builder outputs, and the product builder against the oracle (bit-exact
connectivity, numbering and FP64 rest lengths)."""

import math

import numpy as np
import pytest

import diffloco as dl
from oracle import builders as ob

GOLD = "tests/golden/"


@pytest.mark.parametrize("design,dump", [("robot_2_vertical_and_horizontal.json", "robot_design_data_2.json"),
                                         ("robot_3.json", "robot_design_data_3.json")])
def test_2d_builder_matches_reference_dumps(design, dump):
    objects, springs = ob.MassSpring2D(ob.load_json(GOLD + design)).build()
    ref = ob.load_json(GOLD + dump)
    assert len(objects) == len(ref["objects"]) and len(springs) == len(ref["springs"])
    for o, r in zip(objects, ref["objects"]):
        assert o[0] == r["x"] and o[1] == r["y"]                       # bit-exact
    for s, r in zip(springs, ref["springs"]):
        assert [s[0], s[1], s[2], s[3], s[4]] == [r["vertex_1"], r["vertex_2"], r["length"],
                                                   r["stiffness"], r["actuation"]]


@pytest.mark.parametrize("rid,n,s,act", [(100, 48, 272, 48), (101, 40, 222, 44)])
def test_reference_robot_counts(rid, n, s, act):
    objects, springs, faces = ob.MassSpring3D().build(rid)
    assert (len(objects), len(springs)) == (n, s)
    assert sum(1 for x in springs if x[4] > 0) == act
    assert len(faces) == 12 * (13 if rid == 100 else 10)


def _invariants(objects, springs):
    pairs = set()
    for a, b, L, k, act in springs:
        key = (min(a, b), max(a, b))
        assert key not in pairs
        pairs.add(key)
        assert any(abs(L - 0.05 * math.sqrt(m)) < 1e-15 for m in (1, 2, 3))
    used = {a for a, *_ in springs} | {b for _, b, *_ in springs}
    assert used == set(range(len(objects)))


@pytest.mark.parametrize("rid", [100, 101, 102, 104])
def test_product_builder_equals_oracle_reference_robots(rid):
    o_obj, o_spr, o_faces = ob.MassSpring3D().build(rid)
    d = dl.build_reference(rid)
    assert d.objects.tolist() == [list(p) for p in o_obj]
    assert d.spring_a.tolist() == [s[0] for s in o_spr]
    assert d.spring_b.tolist() == [s[1] for s in o_spr]
    assert d.L0.tolist() == [s[2] for s in o_spr]
    assert d.stiffness.tolist() == [s[3] for s in o_spr]
    assert d.coeff.tolist() == [s[4] for s in o_spr]
    assert d.act_index.tolist() == list(range(len(o_spr)))
    assert [tuple(f) for f in d.faces] == [tuple(f) for f in o_faces]
    _invariants(o_obj, o_spr)


def test_quadruped_design():
    o_obj, o_spr, o_grp = ob.build_quadruped()
    d = dl.build_quadruped()
    assert (d.n_vert, d.n_spring, d.n_act) == (54, 323, 4)
    assert d.objects.tolist() == [list(p) for p in o_obj]
    assert d.L0.tolist() == [s[2] for s in o_spr]
    assert d.act_index.tolist() == o_grp
    assert int((d.coeff > 0).sum()) == 80
    _invariants(o_obj, o_spr)
    # actuated springs belong to leg cubes only (y below the body)
    ys = d.objects[:, 1]
    for a, b, c in zip(d.spring_a, d.spring_b, d.coeff):
        if c > 0:
            assert min(ys[a], ys[b]) < 0.1 + 2 * 0.05 - 1e-12


def test_vertex_and_actuator_csr():
    d = dl.build_reference(100)
    ptr, idx = d.vertex_csr()
    assert ptr[0] == 0 and ptr[-1] == 2 * d.n_spring
    for i in range(d.n_vert):
        ent = idx[ptr[i]:ptr[i + 1]]
        assert np.all(np.diff(ent >> 1) > 0)                          # increasing spring id
        for e in ent:
            s, side = e >> 1, e & 1
            assert (d.spring_b if side else d.spring_a)[s] == i
    q = dl.build_quadruped()
    ap, ax = q.actuator_csr()
    assert ap[-1] == int((q.coeff != 0).sum())
    for o in range(q.n_act):
        assert np.all(q.act_index[ax[ap[o]:ap[o + 1]]] == o)


def test_mpm_2d_builder_counts():
    pos, act, n_act = ob.MPM2D(ob.load_json(GOLD + "robot_mpm_1.json")).build()
    assert len(pos) == 320 and n_act == 4
    assert sorted(set(act)) == [-1, 0, 1, 2, 3]
    assert sum(1 for a in act if a == -1) == 12 * 16


def test_config_derivation_matches_reference_formulae():
    cfg = dl.ConfigSim.from_file(GOLD + "sim_quad.json").get_config()
    assert cfg["robot"]["n_objects"] == 48 and cfg["robot"]["n_springs"] == 272
    assert cfg["nn"]["n_input_states"] == 100 + 3 * 2 * 48 + 150 * 2 + 150
    assert cfg["process"]["dt"] == 0.004 and cfg["simulator"]["gravity"] == -1.8
    assert cfg["simulator"]["dashpot_damping"] == 0.1
    assert abs(cfg["process"]["spring_omega"] - 2 * math.pi / 0.004 / 100) < 1e-12
    q = dl.quadruped_config(batch_size=8).get_config()
    assert q["nn"]["n_input_states"] == 179 and q["robot"]["n_actuators"] == 4


def _fixtures():
    return ob.load_json(GOLD + "builder3d_fixtures.json")


def _assert_same(case, objects, springs, faces):
    assert [list(p) for p in objects] == case["objects"]                  # bit for bit
    assert [[int(s[0]), int(s[1]), float(s[2]), float(s[3]), float(s[4])] for s in springs] \
        == case["springs"]
    assert [list(f) for f in faces] == case["faces"]


FIXTURE_CASES = ["robot_100", "robot_101", "robot_102", "robot_103", "robot_104", "quadruped",
                 "random_12", "random_60", "random_230"]


@pytest.mark.parametrize("name", FIXTURE_CASES)
def test_3d_builders_match_reference_builder_output(name):
    """Oracle and product 3D builders against the reference builder's own
    output (vertex numbering, 28-spring template order, dedupe, FP64 rest
    lengths, faces), bit for bit."""
    case = _fixtures()[name]
    skel = ob.load_json(GOLD + "skeleton.json")
    if case["cubes"] is None:
        rid = int(name.split("_")[1])
        o = ob.MassSpring3D(case["stiffness"], case["actuation"]).build(rid, skeleton=skel)
        _assert_same(case, *o)
        d = dl.build_reference(rid, case["stiffness"], case["actuation"],
                               skeleton=GOLD + "skeleton.json" if rid == 103 else None)
        springs = [[a, b, L, k, c] for a, b, L, k, c in zip(d.spring_a, d.spring_b, d.L0,
                                                           d.stiffness, d.coeff)]
        _assert_same(case, d.objects.tolist(), springs, d.faces)
        assert d.act_index.tolist() == list(range(d.n_spring))
    else:
        cubes = [(tuple(c), act, g) for c, act, g in case["cubes"]]
        b = ob.MassSpring3D(case["stiffness"], case["actuation"])
        for ijk, act, g in cubes:
            b.add_mesh_square(*ijk, actuation=case["actuation"] if act else 0.0, group=g)
        _assert_same(case, b.objects, b.springs, b.faces)
        d = dl.build_voxels(cubes, case["stiffness"], case["actuation"])
        springs = [[a, b_, L, k, c] for a, b_, L, k, c in zip(d.spring_a, d.spring_b, d.L0,
                                                            d.stiffness, d.coeff)]
        _assert_same(case, d.objects.tolist(), springs, d.faces)
        # the group of each spring is its first writer's cube group
        assert d.act_index.tolist() == b.groups


def test_fixture_template_swap_is_detected():
    """A plausible builder mistake -- two entries of the 28-spring template
    swapped -- changes the spring order the fixture pins."""
    case = _fixtures()["quadruped"]
    b = ob.MassSpring3D(case["stiffness"], case["actuation"])
    orig = ob.MassSpring3D.add_mesh_square

    def swapped(self, i, j, k, actuation=0.0, group=0):
        a, bb, c, d, e, f, g, h = self._cube_points(i, j, k)
        s, A = self.spring_stiffness, actuation
        order = [(a, bb, A), (c, d, A), (e, f, A), (g, h, A),
                 (a, c, 0), (bb, d, 0), (f, h, 0), (e, g, 0),      # first two x-edges swapped
                 (bb, f, 0), (d, h, 0), (a, e, 0), (c, g, 0),
                 (bb, g, 0), (d, e, 0), (f, c, 0), (h, a, 0),
                 (e, c, 0), (a, g, 0), (h, bb, 0), (d, f, 0),
                 (e, bb, A), (a, f, A), (h, c, A), (d, g, A),
                 (f, g, A), (e, h, A), (a, d, A), (c, bb, A)]
        for p, q, act in order:
            self.add_mesh_spring(p, q, s, act, group)
        self._faces(a, bb, c, d, e, f, g, h)

    for (ijk, act, g) in [(tuple(c), a, g) for c, a, g in case["cubes"]]:
        swapped(b, *ijk, actuation=case["actuation"] if act else 0.0, group=g)
    got = [[int(s[0]), int(s[1]), float(s[2]), float(s[3]), float(s[4])] for s in b.springs]
    assert got != case["springs"]
    assert orig is ob.MassSpring3D.add_mesh_square


def test_robot_103_from_config_file(tmp_path):
    """sim_spider.json selects robot 103; the loader reads skeleton.json next
    to the experiment file (the reference reads cfg3d/skeleton.json)."""
    import shutil
    for f in ("sim_spider.json", "skeleton.json"):
        shutil.copy(GOLD + f, tmp_path / f)
    cfg = dl.ConfigSim.from_file(str(tmp_path / "sim_spider.json")).get_config()
    case = _fixtures()["robot_103"]
    assert cfg["robot"]["n_objects"] == len(case["objects"]) == 39
    assert cfg["robot"]["n_springs"] == len(case["springs"]) == 175
    assert [list(p) for p in cfg["robot"]["objects"]] == case["objects"]
