"""Restated robot builders (TEST INFRASTRUCTURE ONLY -- see oracle/__init__.py).

Follows ``multitask/robot_design.py`` literally, including its list-membership
dedupe (first writer wins) and its FP64 ``**0.5`` rest lengths, so that vertex
numbering, spring order and rest lengths are bit-exact with the reference.

Pinned by: ``tests/golden/robot_design_data_{2,3}.json`` (the reference's own
dumped builder outputs for the 2D builder), ``tests/golden/builder3d_fixtures.json``
(outputs of the reference's own 3D builder for robots 100-104, the quadruped
cube list and three random voxel robots, written by
``tests/golden/make_builder_fixtures.py``), spring/vertex counts of robot 100
(``robot_design.py:429-443``), and lattice invariants (rest lengths in
{0.05, 0.05*sqrt2, 0.05*sqrt3}, no duplicate pairs, contiguous ids).
"""

import json


class MassSpring2D:
    """``RobotDesignMassSpring`` (``robot_design.py:113-317``)."""

    def __init__(self, cfg):
        self.config = cfg
        self.objects = []
        self.springs = []
        self.points = []
        self.point_id = []
        self.mesh_springs = []

    def add_object(self, x):                       # robot_design.py:236-238
        self.objects.append(x)
        return len(self.objects) - 1

    def add_spring(self, a, b, length=None, stiffness=1, actuation=0.1):  # :240-244
        if length is None:
            length = ((self.objects[a][0] - self.objects[b][0]) ** 2 +
                      (self.objects[a][1] - self.objects[b][1]) ** 2) ** 0.5
        self.springs.append([a, b, length, stiffness, actuation])

    def add_mesh_point(self, i, j):                # :246-251
        if (i, j) not in self.points:
            pid = self.add_object((i * 0.05 + 0.1, j * 0.05 + 0.1))
            self.points.append((i, j))
            self.point_id.append(pid)
        return self.point_id[self.points.index((i, j))]

    def add_mesh_spring(self, a, b, s, act):       # :253-258
        if (a, b) in self.mesh_springs or (b, a) in self.mesh_springs:
            return
        self.mesh_springs.append((a, b))
        self.add_spring(a, b, stiffness=s, actuation=act)

    def add_mesh_square(self, i, j, stiffness=3e4, actuation=0.0, active_spring=None):  # :260-273
        a = self.add_mesh_point(i, j)
        b = self.add_mesh_point(i, j + 1)
        c = self.add_mesh_point(i + 1, j)
        d = self.add_mesh_point(i + 1, j + 1)
        if active_spring is None:
            active_spring = [1] * 6
        links = [[a, b], [a, c], [a, d], [b, c], [b, d], [c, d]]
        for n, (p, q) in enumerate(links):
            self.add_mesh_spring(p, q, stiffness, actuation * active_spring[n])

    def build(self):                               # :129-175 (square meshes only)
        d = self.config["design"]
        for n, anchor in enumerate(d["anchor"]):
            mesh = d["mesh_type"][n]
            if d["actuation_enable"][n] == 1:
                active = self.config["active_spring_template"][mesh][d["active_spring_mechanism"][n]]
            else:
                active = self.config["active_spring_template"][mesh]["dummy"]
            if mesh != "square":
                raise NotImplementedError(mesh)
            self.add_mesh_square(anchor[0], anchor[1],
                                 stiffness=d["physical_parameter"]["stiffness"][n],
                                 actuation=d["physical_parameter"]["actuation"][n],
                                 active_spring=active)
        return self.objects, self.springs


class MassSpring3D:
    """``RobotDesignMassSpring3D`` (``robot_design.py:319-591``).

    ``add_mesh_square`` takes an extra ``group`` (actuator group id recorded
    per spring, first writer wins like the spring itself); the reference's
    springs carry no group because its MLP drives every spring separately.
    """

    def __init__(self, spring_stiffness=1e4, spring_actuation=0.25):
        self.spring_stiffness = spring_stiffness
        self.spring_actuation = spring_actuation
        self.objects = []
        self.springs = []
        self.groups = []
        self.points = []
        self.point_id = []
        self.mesh_springs = []
        self.faces = []

    def add_object(self, x):                       # :340-342
        self.objects.append(x)
        return len(self.objects) - 1

    def add_spring(self, a, b, length=None, stiffness=1, actuation=0.1, group=0):  # :344-349
        if length is None:
            length = ((self.objects[a][0] - self.objects[b][0]) ** 2 +
                      (self.objects[a][1] - self.objects[b][1]) ** 2 +
                      (self.objects[a][2] - self.objects[b][2]) ** 2) ** 0.5
        self.springs.append([a, b, length, stiffness, actuation])
        self.groups.append(group)

    def add_mesh_point(self, i, j, k):             # :351-356
        if (i, j, k) not in self.points:
            pid = self.add_object((i * 0.05 + 0.1, j * 0.05 + 0.1, k * 0.05 + 0.1))
            self.points.append((i, j, k))
            self.point_id.append(pid)
        return self.point_id[self.points.index((i, j, k))]

    def add_mesh_spring(self, a, b, s, act, group=0):  # :365-370
        if (a, b) in self.mesh_springs or (b, a) in self.mesh_springs:
            return
        self.mesh_springs.append((a, b))
        self.add_spring(a, b, stiffness=s, actuation=act, group=group)

    def _cube_points(self, i, j, k):
        P = self.add_mesh_point
        return (P(i, j, k), P(i, j + 1, k), P(i + 1, j, k), P(i + 1, j + 1, k),
                P(i, j, k + 1), P(i, j + 1, k + 1), P(i + 1, j, k + 1), P(i + 1, j + 1, k + 1))

    def _faces(self, a, b, c, d, e, f, g, h):     # :419-427
        def sq(p, q, r, s):
            self.faces.append((p, q, r))
            self.faces.append((s, p, r))
        sq(a, b, d, c); sq(b, f, h, d); sq(f, e, g, h)
        sq(e, a, c, g); sq(h, g, c, d); sq(b, a, e, f)

    def add_mesh_square(self, i, j, k, actuation=0.0, group=0):  # :372-427
        a, b, c, d, e, f, g, h = self._cube_points(i, j, k)
        s, A = self.spring_stiffness, actuation
        order = [(a, b, A), (c, d, A), (e, f, A), (g, h, A),
                 (b, d, 0), (a, c, 0), (f, h, 0), (e, g, 0),
                 (b, f, 0), (d, h, 0), (a, e, 0), (c, g, 0),
                 (b, g, 0), (d, e, 0), (f, c, 0), (h, a, 0),
                 (e, c, 0), (a, g, 0), (h, b, 0), (d, f, 0),
                 (e, b, A), (a, f, A), (h, c, A), (d, g, A),
                 (f, g, A), (e, h, A), (a, d, A), (c, b, A)]
        for p, q, act in order:
            self.add_mesh_spring(p, q, s, act, group)
        self._faces(a, b, c, d, e, f, g, h)

    def add_mesh_lying(self, i, j, k, actuation=0.0, group=0):  # :507-562
        a, b, c, d, e, f, g, h = self._cube_points(i, j, k)
        s, A = self.spring_stiffness, actuation
        order = [(a, b, 0), (c, d, 0), (e, f, 0), (g, h, 0),
                 (b, d, 0), (a, c, 0), (f, h, 0), (e, g, 0),
                 (b, f, A), (d, h, A), (a, e, A), (c, g, A),
                 (b, g, A), (d, e, A), (f, c, A), (h, a, A),
                 (e, c, A), (a, g, A), (h, b, A), (d, f, A),
                 (e, b, A), (a, f, A), (h, c, A), (d, g, A),
                 (f, g, 0), (e, h, 0), (a, d, 0), (c, b, 0)]
        for p, q, act in order:
            self.add_mesh_spring(p, q, s, act, group)
        self._faces(a, b, c, d, e, f, g, h)

    # robots 100..104 (:429-570); robotD reads the skeleton graph (cfg3d/skeleton.json)
    def robotA(self):
        for ijk in [(0, 0, 0), (2, 0, 2), (0, 0, 2), (2, 0, 0)]:
            self.add_mesh_square(*ijk, actuation=self.spring_actuation)
        for ijk in [(0, 1, 0), (0, 1, 1), (0, 1, 2), (1, 1, 0), (1, 1, 1),
                    (1, 1, 2), (2, 1, 0), (2, 1, 1), (2, 1, 2)]:
            self.add_mesh_square(*ijk, actuation=0)

    def robotB(self):
        for ijk in [(0, 0, 1), (1, 0, 0), (1, 0, 2), (2, 0, 1)]:
            self.add_mesh_square(*ijk, actuation=self.spring_actuation)
        for ijk in [(0, 1, 1), (1, 1, 0), (1, 1, 2), (2, 1, 1), (1, 1, 1), (1, 2, 1)]:
            self.add_mesh_square(*ijk, actuation=0)

    def robotC(self):
        for ijk in [(0, 0, 0), (2, 0, 0), (1, 0, 2)]:
            self.add_mesh_square(*ijk, actuation=self.spring_actuation)
        for ijk in [(0, 1, 0), (2, 1, 0), (1, 1, 2), (0, 1, 1), (1, 1, 1), (2, 1, 1)]:
            self.add_mesh_square(*ijk, actuation=0)

    def robotE(self):
        for kk in range(6):
            self.add_mesh_lying(0, 0, kk, actuation=self.spring_actuation)

    def robotD(self, skeleton):                    # :458-505, skeleton = parsed JSON
        for v in skeleton["nodes"]:
            pos = list(skeleton["nodes"][v])
            pos[0] = pos[0] * 0.05
            pos[1] = pos[1] * 0.05 + 0.212
            pos[2] = pos[2] * 0.05 * 0.75
            self.objects.append(pos)
        motors = {(9, 21), (36, 24), (29, 26), (14, 11), (14, 26), (10, 9), (29, 11), (25, 24),
                  (25, 26), (36, 9), (21, 24), (10, 11)}
        s = set()
        for e in skeleton["links"]:
            a, b = e
            s.add((a, b))
            s.add((b, a))
            act = self.spring_actuation if (a, b) in motors or (b, a) in motors else 0.
            self.add_mesh_spring(a, b, self.spring_stiffness, act)
        for a, b in list(motors) + [(14, 29), (29, 36), (36, 21), (21, 14), (11, 5), (26, 5),
                                    (24, 35), (9, 20), (26, 35), (11, 20), (9, 8), (24, 8),
                                    (11, 31), (26, 16), (24, 23), (9, 38)]:
            act = self.spring_actuation if (a, b) in motors or (b, a) in motors else 0.
            self.add_mesh_spring(a, b, self.spring_stiffness, act)
        n = len(self.objects)
        for a in range(n):
            for b in range(n):
                for c in range(n):
                    if a < b < c and (a, b) in s and (b, c) in s and (a, c) in s:
                        self.faces.append((a, b, c))

    def build(self, robot_id, skeleton=None):     # :572-587
        if robot_id == 103:
            self.robotD(skeleton)
        else:
            {100: self.robotA, 101: self.robotB, 102: self.robotC, 104: self.robotE}[robot_id]()
        return self.objects, self.springs, self.faces


def quadruped_cubes():
    """Cube list of the 3D quadruped workload: a 4x1x2 body (y = 2) over four
    1x2x1 legs (y = 0, 1) under the body's corner cubes.  Legs come first, in
    leg order, each leg bottom-up; leg ``g`` is actuator group ``g``."""
    cubes = []
    for g, (i, k) in enumerate([(0, 0), (3, 0), (0, 1), (3, 1)]):
        for j in (0, 1):
            cubes.append(((i, j, k), True, g))
    for i in range(4):
        for k in range(2):
            cubes.append(((i, 2, k), False, 0))
    return cubes


def build_quadruped(stiffness=3e4, actuation=0.1):
    b = MassSpring3D(spring_stiffness=stiffness, spring_actuation=actuation)
    for ijk, act, g in quadruped_cubes():
        b.add_mesh_square(*ijk, actuation=actuation if act else 0.0, group=g)
    return b.objects, b.springs, b.groups


def build_voxels(cubes, stiffness=3e4, actuation=0.1):
    """Generic cube list ``[((i,j,k), actuated, group), ...]`` through the
    restated ``add_mesh_square``."""
    b = MassSpring3D(spring_stiffness=stiffness, spring_actuation=actuation)
    for ijk, act, g in cubes:
        b.add_mesh_square(*ijk, actuation=actuation if act else 0.0, group=g)
    return b.objects, b.springs, b.groups


class MPM2D:
    """``RobotDesignMPM`` (``robot_design.py:38-110``)."""

    def __init__(self, cfg):
        self.config = cfg
        p = cfg["design_parameter"]
        self.square_size = p["square_size"]
        self.n_num = p["particle_in_square_edge"]
        self.actuator_num = p["actuator_num"]
        off = p.get("offset", {"x": 0, "y": 0})
        self.offset = (off["x"], off["y"])
        self.pos = []
        self.actuator_id = []

    def add_particle(self, pos, act_id):           # :65-70
        assert -1 <= act_id < self.actuator_num
        self.pos.append((self.offset[0] + pos[0], self.offset[1] + pos[1]))
        self.actuator_id.append(act_id)

    def add_square(self, pos, act_id):             # :72-79
        dx = self.square_size / self.n_num
        for i in range(self.n_num):
            for j in range(self.n_num):
                self.add_particle((pos[0] * self.square_size + (i + 0.5) * dx,
                                   pos[1] * self.square_size + (j + 0.5) * dx), act_id)

    def build(self):                               # :81-103
        d = self.config["design"]
        for n, mesh in enumerate(d["mesh_type"]):
            assert mesh == "rectangle"
            pos, shape = d["anchor"][n], d["shape"][n]
            for i in range(shape[0]):
                for j in range(shape[1]):
                    self.add_square((pos[0] + i, pos[1] + j), d["actuator"][n])
        return self.pos, self.actuator_id, self.actuator_num


def load_json(path):
    with open(path) as fh:
        return json.load(fh)
