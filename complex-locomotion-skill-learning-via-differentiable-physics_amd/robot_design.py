"""Voxel mass-spring robot builder and the device topology it produces.

Same numbering rules as the reference builder
(``multitask/robot_design.py:340-427`` and ``:507-591``): lattice vertices get
ids in first-insertion order, each cube emits its 28 springs in a fixed
template order, a spring whose unordered endpoint pair already exists is
dropped (first writer wins), rest lengths are the FP64 ``**0.5`` distance of
the endpoints.  Dedupe uses hash maps instead of list scans, so a 400-cube
robot builds in milliseconds.

Each spring also records the MLP output that drives it (``act_index``): the
reference drives every spring by its own output (identity map); voxel robots
with actuator groups map every actuated spring of a cube to the cube's group.
"""

import math
from dataclasses import dataclass

import numpy as np

EDGE = 0.05
ORIGIN = 0.1

# Cube corners in template order: (di, dj, dk) for a..h
_CORNERS = ((0, 0, 0), (0, 1, 0), (1, 0, 0), (1, 1, 0),
            (0, 0, 1), (0, 1, 1), (1, 0, 1), (1, 1, 1))
_a, _b, _c, _d, _e, _f, _g, _h = range(8)
# (corner, corner, actuated?) -- upright cube (robot_design.py:384-417):
# y-edges, x-edges, z-edges, body diagonals, horizontal-face diagonals,
# x-face diagonals, z-face diagonals; y-edges and vertical faces actuated.
UPRIGHT = ((_a, _b, 1), (_c, _d, 1), (_e, _f, 1), (_g, _h, 1),
           (_b, _d, 0), (_a, _c, 0), (_f, _h, 0), (_e, _g, 0),
           (_b, _f, 0), (_d, _h, 0), (_a, _e, 0), (_c, _g, 0),
           (_b, _g, 0), (_d, _e, 0), (_f, _c, 0), (_h, _a, 0),
           (_e, _c, 0), (_a, _g, 0), (_h, _b, 0), (_d, _f, 0),
           (_e, _b, 1), (_a, _f, 1), (_h, _c, 1), (_d, _g, 1),
           (_f, _g, 1), (_e, _h, 1), (_a, _d, 1), (_c, _b, 1))
# lying cube (robot_design.py:519-552): z-edges, diagonals actuated
LYING = ((_a, _b, 0), (_c, _d, 0), (_e, _f, 0), (_g, _h, 0),
         (_b, _d, 0), (_a, _c, 0), (_f, _h, 0), (_e, _g, 0),
         (_b, _f, 1), (_d, _h, 1), (_a, _e, 1), (_c, _g, 1),
         (_b, _g, 1), (_d, _e, 1), (_f, _c, 1), (_h, _a, 1),
         (_e, _c, 1), (_a, _g, 1), (_h, _b, 1), (_d, _f, 1),
         (_e, _b, 1), (_a, _f, 1), (_h, _c, 1), (_d, _g, 1),
         (_f, _g, 0), (_e, _h, 0), (_a, _d, 0), (_c, _b, 0))
# two triangles per face (robot_design.py:419-427)
_FACES = ((_a, _b, _d, _c), (_b, _f, _h, _d), (_f, _e, _g, _h),
          (_e, _a, _c, _g), (_h, _g, _c, _d), (_b, _a, _e, _f))


class VoxelRobot3D:
    """Incremental 3D voxel builder (RobotDesignMassSpring3D equivalent)."""

    def __init__(self, spring_stiffness=1e4, spring_actuation=0.25):
        self.spring_stiffness = spring_stiffness
        self.spring_actuation = spring_actuation
        self.objects = []
        self.springs = []     # [a, b, L0, stiffness, actuation]
        self.groups = []
        self.faces = []
        self._ids = {}
        self._pairs = set()

    def point(self, i, j, k):
        key = (i, j, k)
        pid = self._ids.get(key)
        if pid is None:
            pid = len(self.objects)
            self._ids[key] = pid
            self.objects.append((i * EDGE + ORIGIN, j * EDGE + ORIGIN, k * EDGE + ORIGIN))
        return pid

    def spring(self, a, b, stiffness, actuation, group=0):
        if (a, b) in self._pairs or (b, a) in self._pairs:
            return
        self._pairs.add((a, b))
        pa, pb = self.objects[a], self.objects[b]
        length = ((pa[0] - pb[0]) ** 2 + (pa[1] - pb[1]) ** 2 + (pa[2] - pb[2]) ** 2) ** 0.5
        self.springs.append([a, b, length, stiffness, actuation])
        self.groups.append(group)

    def cube(self, i, j, k, actuation=0.0, group=0, template=UPRIGHT):
        ids = [self.point(i + di, j + dj, k + dk) for di, dj, dk in _CORNERS]
        for p, q, act in template:
            self.spring(ids[p], ids[q], self.spring_stiffness, actuation if act else 0, group)
        for p, q, r, s in _FACES:
            self.faces.append((ids[p], ids[q], ids[r]))
            self.faces.append((ids[s], ids[p], ids[r]))

    # reference robots (robot_design.py:429-570); robot 103 reads a skeleton file
    _REF = {
        100: ([(0, 0, 0), (2, 0, 2), (0, 0, 2), (2, 0, 0)],
              [(0, 1, 0), (0, 1, 1), (0, 1, 2), (1, 1, 0), (1, 1, 1), (1, 1, 2), (2, 1, 0),
               (2, 1, 1), (2, 1, 2)]),
        101: ([(0, 0, 1), (1, 0, 0), (1, 0, 2), (2, 0, 1)],
              [(0, 1, 1), (1, 1, 0), (1, 1, 2), (2, 1, 1), (1, 1, 1), (1, 2, 1)]),
        102: ([(0, 0, 0), (2, 0, 0), (1, 0, 2)],
              [(0, 1, 0), (2, 1, 0), (1, 1, 2), (0, 1, 1), (1, 1, 1), (2, 1, 1)]),
    }

    # Robot 103 (robot_design.py:458-505): a skeleton graph read from a JSON
    # file ({"nodes": {id: [x, y, z]}, "links": [[a, b], ...]}, the
    # reference's cfg3d/skeleton.json).  Motor springs are added after the
    # links, in the order the reference's motor set iterates; then the extra
    # passive braces.  Faces are the link graph's triangles a < b < c.
    SKELETON_MOTORS = ((36, 24), (10, 11), (29, 26), (14, 26), (9, 21), (25, 24), (36, 9),
                       (21, 24), (10, 9), (29, 11), (25, 26), (14, 11))
    SKELETON_BRACES = ((14, 29), (29, 36), (36, 21), (21, 14), (11, 5), (26, 5), (24, 35),
                       (9, 20), (26, 35), (11, 20), (9, 8), (24, 8), (11, 31), (26, 16),
                       (24, 23), (9, 38))

    def skeleton(self, data):
        for pos in data["nodes"].values():
            self.objects.append((pos[0] * 0.05, pos[1] * 0.05 + 0.212, pos[2] * 0.05 * 0.75))
        motors = set(self.SKELETON_MOTORS)
        motor = lambda a, b: (a, b) in motors or (b, a) in motors
        linked = set()
        for a, b in data["links"]:
            linked.add((a, b))
            linked.add((b, a))
            self.spring(a, b, self.spring_stiffness, self.spring_actuation if motor(a, b) else 0.0)
        for a, b in self.SKELETON_MOTORS + self.SKELETON_BRACES:
            self.spring(a, b, self.spring_stiffness, self.spring_actuation if motor(a, b) else 0.0)
        n = len(self.objects)
        adj = [sorted(b for b in range(n) if (a, b) in linked) for a in range(n)]
        for a in range(n):
            for b in adj[a]:
                if b <= a:
                    continue
                for c in adj[b]:
                    if c > b and (a, c) in linked:
                        self.faces.append((a, b, c))
        return self.objects, self.springs, self.faces

    def build_reference(self, robot_id, skeleton=None):
        """Reference robots 100-104.  Robot 103 needs ``skeleton``: the
        skeleton graph as a dict or the path of its JSON file."""
        if robot_id == 103:
            if skeleton is None:
                raise ValueError("robot 103 is built from a skeleton file (cfg3d/skeleton.json)")
            if not isinstance(skeleton, dict):
                import json
                with open(skeleton) as fh:
                    skeleton = json.load(fh)
            return self.skeleton(skeleton)
        if robot_id == 104:
            for kk in range(6):
                self.cube(0, 0, kk, actuation=self.spring_actuation, template=LYING)
        elif robot_id in self._REF:
            legs, body = self._REF[robot_id]
            for ijk in legs:
                self.cube(*ijk, actuation=self.spring_actuation)
            for ijk in body:
                self.cube(*ijk, actuation=0)
        else:
            raise ValueError(f"robot id {robot_id} is not a voxel robot")
        return self.objects, self.springs, self.faces


def quadruped_cubes():
    """The quadruped workload: a 4x1x2 body of cubes (y = 2) over four 1x2x1
    legs (y = 0, 1) under the body's corner cubes.  Legs first, each bottom-up;
    leg g is actuator group g.  Entries: ((i, j, k), actuated, group)."""
    cubes = []
    for g, (i, k) in enumerate(((0, 0), (3, 0), (0, 1), (3, 1))):
        cubes += [((i, 0, k), True, g), ((i, 1, k), True, g)]
    cubes += [((i, 2, k), False, 0) for i in range(4) for k in range(2)]
    return cubes


@dataclass
class Design:
    """Host-side topology of one robot design (numpy)."""
    objects: np.ndarray        # [N,3] f64
    spring_a: np.ndarray       # [S] int32
    spring_b: np.ndarray
    L0: np.ndarray             # [S] f64
    stiffness: np.ndarray      # [S] f64 as built (solver applies its scale)
    coeff: np.ndarray          # [S] f64 actuation coefficient
    act_index: np.ndarray      # [S] int32
    n_act: int
    faces: list = None

    @property
    def n_vert(self):
        return self.objects.shape[0]

    @property
    def n_spring(self):
        return self.spring_a.shape[0]

    def vertex_csr(self):
        """vertex -> incident springs, entries 2*s + side, increasing s."""
        S, N = self.n_spring, self.n_vert
        vert = np.concatenate([self.spring_a, self.spring_b]).astype(np.int64)
        ent = np.concatenate([2 * np.arange(S), 2 * np.arange(S) + 1])
        order = np.lexsort((ent, vert))
        ptr = np.zeros(N + 1, np.int64)
        np.add.at(ptr, vert + 1, 1)
        return np.cumsum(ptr).astype(np.int32), ent[order].astype(np.int32)

    def actuator_csr(self):
        """MLP output -> driven springs (coefficient != 0), increasing s."""
        live = np.nonzero(self.coeff != 0)[0]
        outs = self.act_index[live].astype(np.int64)
        order = np.lexsort((live, outs))
        ptr = np.zeros(self.n_act + 1, np.int64)
        np.add.at(ptr, outs + 1, 1)
        return np.cumsum(ptr).astype(np.int32), live[order].astype(np.int32)


def make_design(objects, springs, act_index=None, n_act=None, faces=None):
    S = len(springs)
    if act_index is None:
        act_index = np.arange(S)
        n_act = S
    act_index = np.asarray(act_index, np.int32)
    if n_act is None:
        n_act = int(act_index.max()) + 1 if S else 1
    sp = springs
    return Design(objects=np.asarray(objects, np.float64).reshape(-1, 3),
                  spring_a=np.array([s[0] for s in sp], np.int32),
                  spring_b=np.array([s[1] for s in sp], np.int32),
                  L0=np.array([s[2] for s in sp], np.float64),
                  stiffness=np.array([s[3] for s in sp], np.float64),
                  coeff=np.array([s[4] for s in sp], np.float64),
                  act_index=act_index, n_act=int(n_act), faces=faces)


def build_voxels(cubes, stiffness=3e4, actuation=0.1, n_act=None):
    """Build a grouped voxel robot from ``[((i,j,k), actuated, group), ...]``."""
    b = VoxelRobot3D(spring_stiffness=stiffness, spring_actuation=actuation)
    for ijk, act, g in cubes:
        b.cube(*ijk, actuation=actuation if act else 0.0, group=g)
    if n_act is None:
        n_act = max(g for _, _, g in cubes) + 1
    return make_design(b.objects, b.springs, act_index=b.groups, n_act=n_act, faces=b.faces)


def build_quadruped(stiffness=3e4, actuation=0.1):
    return build_voxels(quadruped_cubes(), stiffness, actuation, n_act=4)


def build_reference(robot_id, stiffness=1e4, actuation=0.25, skeleton=None):
    """Reference robot with one MLP output per spring (diffphy_trainer.py:178-185);
    robot 103 takes its skeleton graph (dict or JSON path)."""
    b = VoxelRobot3D(spring_stiffness=stiffness, spring_actuation=actuation)
    objects, springs, faces = b.build_reference(robot_id, skeleton=skeleton)
    return make_design(objects, springs, faces=faces)


_DIRS = ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1))


def random_voxel_cubes(rng, n_cubes, n_groups):
    """Connected random accretion: start from one cube and repeatedly attach
    a new cube to a face of a uniformly chosen existing cube (never below
    y = 0).  Actuator groups are spatial clusters: ``n_groups`` seed cubes,
    every cube joins its nearest seed (ties to the lower seed index)."""
    cells = [(0, 0, 0)]
    have = {(0, 0, 0)}
    while len(cells) < n_cubes:
        b = cells[int(rng.integers(len(cells)))]
        d = _DIRS[int(rng.integers(6))]
        c = (b[0] + d[0], b[1] + d[1], b[2] + d[2])
        if c[1] < 0 or c in have:
            continue
        have.add(c)
        cells.append(c)
    g = min(n_groups, n_cubes)
    seeds = np.array([cells[i] for i in rng.choice(n_cubes, size=g, replace=False)], np.float64)
    pos = np.array(cells, np.float64)
    dist = ((pos[:, None, :] - seeds[None]) ** 2).sum(-1)
    group = dist.argmin(1)
    return [(c, True, int(k)) for c, k in zip(cells, group)], g


def random_voxel_robot(seed, n_cubes=None, n_groups=None, stiffness=3e4, actuation=0.1):
    """One random robot: ``n_cubes`` log-uniform in [8, 400] and ``n_groups``
    uniform in [4, 64] unless given."""
    rng = np.random.default_rng(seed)
    if n_cubes is None:
        n_cubes = int(round(math.exp(rng.uniform(math.log(8), math.log(400)))))
    if n_groups is None:
        n_groups = int(rng.integers(4, 65))
    cubes, g = random_voxel_cubes(rng, n_cubes, n_groups)
    return build_voxels(cubes, stiffness, actuation, n_act=g)


def random_robot_batch(n, seed, horizon_range=(200, 1000), cube_range=(8, 400), group_range=(4, 64)):
    """``n`` random designs with horizons uniform in ``horizon_range``."""
    rng = np.random.default_rng(seed)
    designs, horizons = [], []
    for _ in range(n):
        nc = int(round(math.exp(rng.uniform(math.log(cube_range[0]), math.log(cube_range[1])))))
        ng = int(rng.integers(group_range[0], group_range[1] + 1))
        cubes, g = random_voxel_cubes(rng, nc, ng)
        designs.append(build_voxels(cubes, n_act=g))
        horizons.append(int(rng.integers(horizon_range[0], horizon_range[1] + 1)))
    return designs, horizons
