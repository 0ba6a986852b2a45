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
