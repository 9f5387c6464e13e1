"""Deterministic scene generators for the rigid-body contact step.

Every scene is a pair of numpy structured arrays (bodies, shapes) in the pxg ABI layouts, plus the
solver settings it is quoted with. Values are float32 exactly as the GPU and the CPU oracle see
them. Inertia helpers restate physxcoopa/dynamics/inertia.h:36-107 in float32 arithmetic.

Scenes:
  * config1_scene: 64 unit boxes on a 4x4x4 grid over a 50x50x1 m ground (the reference's CPU case).
  * config2_scene: the 16,384-body mixed pile (40% boxes, 40% spheres, 20% Z capsules).
  * config3_scenes: thousands of independent small scenes (box pyramids and random drops).
  * small fixtures mirroring physxcoopa/tests/support/physics_fixtures.h.
"""
import math

import numpy as np

from ._abi import (BODY_DTYPE, SHAPE_DTYPE, PXG_STATIC, PXG_DYNAMIC, PXG_SPHERE, PXG_BOX, PXG_CAPSULE, PXG_CONVEX_HULL, PXG_TRIANGLE_MESH, TriangleMesh,
                   PXG_AVERAGE)

F = np.float32
PI_F = F(3.14159265358979323846)
EPS_F = F(1e-6)


def _fmax(a, b):  # std::max(a, b) = (a < b) ? b : a
    return b if a < b else a


def box_inverse_inertia(half_extents, mass):
    """physxcoopa/dynamics/inertia.h:36-47 in float32."""
    mass = F(mass)
    if mass <= F(0):
        return (F(0), F(0), F(0))
    e = [F(2.0) * F(h) for h in half_extents]
    e2 = [x * x for x in e]
    ix = (mass * (e2[1] + e2[2])) / F(12.0)
    iy = (mass * (e2[0] + e2[2])) / F(12.0)
    iz = (mass * (e2[0] + e2[1])) / F(12.0)
    return (F(1.0) / _fmax(ix, EPS_F), F(1.0) / _fmax(iy, EPS_F), F(1.0) / _fmax(iz, EPS_F))


def sphere_inverse_inertia(radius, mass):
    """physxcoopa/dynamics/inertia.h:56-61 in float32."""
    mass, r = F(mass), F(radius)
    if mass <= F(0):
        return (F(0), F(0), F(0))
    i = ((F(0.4) * mass) * r) * r
    inv = F(1.0) / _fmax(i, EPS_F)
    return (inv, inv, inv)


def capsule_inverse_inertia(radius, half_height, mass):
    """physxcoopa/dynamics/inertia.h:87-107 in float32: (perp, perp, axial) with the long axis Z."""
    mass, r, hh = F(mass), F(radius), F(half_height)
    if mass <= F(0):
        return (F(0), F(0), F(0))
    vol_cyl = ((PI_F * r) * r) * (F(2.0) * hh)
    vol_hemi = (((F(4.0 / 3.0) * PI_F) * r) * r) * r
    vol_total = vol_cyl + vol_hemi
    if vol_total < EPS_F:
        return sphere_inverse_inertia(r, mass)
    m_cyl = mass * (vol_cyl / vol_total)
    m_hemi = mass * (vol_hemi / vol_total)
    i_axial = ((F(0.5) * m_cyl) * r) * r + ((F(0.4) * m_hemi) * r) * r
    i_perp = ((m_cyl * ((F(3.0) * r) * r + (F(4.0) * hh) * hh)) / F(12.0) + ((F(0.4) * m_hemi) * r) * r) \
        + (m_hemi * hh) * hh
    return (F(1.0) / _fmax(i_perp, EPS_F), F(1.0) / _fmax(i_perp, EPS_F), F(1.0) / _fmax(i_axial, EPS_F))


def quat_axis_angle(axis, angle):
    """(x, y, z, w) of a rotation of `angle` radians about unit `axis`."""
    s = math.sin(angle * 0.5)
    return (axis[0] * s, axis[1] * s, axis[2] * s, math.cos(angle * 0.5))


def random_quaternion(rng):
    """Uniform random unit quaternion (Shoemake's method), (x, y, z, w)."""
    u1, u2, u3 = rng.random(3)
    a, b = math.sqrt(1.0 - u1), math.sqrt(u1)
    q = (a * math.sin(2 * math.pi * u2), a * math.cos(2 * math.pi * u2),
         b * math.sin(2 * math.pi * u3), b * math.cos(2 * math.pi * u3))
    n = math.sqrt(sum(c * c for c in q))
    return tuple(c / n for c in q)


class SceneBuilder:
    """Collects bodies and one shape per body (or none) into pxg ABI arrays."""

    def __init__(self):
        self._bodies = []
        self._shapes = []

    def add_body(self, btype=PXG_DYNAMIC, position=(0, 0, 0), orientation=(0, 0, 0, 1), mass=1.0,
                 inv_mass=None, inv_inertia=(1, 1, 1), velocity=(0, 0, 0), angular_velocity=(0, 0, 0),
                 linear_drag=0.0, angular_drag=0.05, use_gravity=True, awake=True, constraints=0):
        b = np.zeros((), dtype=BODY_DTYPE)
        b["type"] = btype
        b["alive"] = 1
        b["awake"] = 1 if awake else 0
        b["use_gravity"] = 1 if use_gravity else 0
        b["constraints"] = constraints
        b["position"] = position
        b["orientation"] = orientation
        b["linear_velocity"] = velocity
        b["angular_velocity"] = angular_velocity
        b["mass"] = mass
        b["inv_mass"] = (F(1.0) / F(mass) if F(mass) > 0 else F(0)) if inv_mass is None else inv_mass
        if btype != PXG_DYNAMIC and inv_mass is None:
            b["inv_mass"] = 1.0  # the reference's Body default; never read for non-dynamic bodies
        b["inv_inertia_local"] = inv_inertia
        b["linear_drag"] = linear_drag
        b["angular_drag"] = angular_drag
        b["prev_position"] = position
        b["prev_orientation"] = orientation
        self._bodies.append(b)
        return len(self._bodies) - 1

    def add_shape(self, owner, stype, radius=0.5, half_extents=(0.5, 0.5, 0.5), capsule_radius=0.5,
                  capsule_half_height=0.5, capsule_axis=2, friction=0.6, restitution=0.0,
                  friction_combine=PXG_AVERAGE, restitution_combine=PXG_AVERAGE, is_trigger=False, layer=0,
                  local_center=(0, 0, 0), local_rotation=(0, 0, 0, 1)):
        s = np.zeros((), dtype=SHAPE_DTYPE)
        s["owner"] = owner
        s["type"] = stype
        s["enabled"] = 1
        s["is_trigger"] = 1 if is_trigger else 0
        s["layer"] = layer
        s["capsule_axis"] = capsule_axis
        s["friction_combine"] = friction_combine
        s["restitution_combine"] = restitution_combine
        s["local_center"] = local_center
        s["radius"] = radius
        s["local_rotation"] = local_rotation
        s["half_extents"] = half_extents
        s["capsule_radius"] = capsule_radius
        s["capsule_half_height"] = capsule_half_height
        s["dynamic_friction"] = friction
        s["static_friction"] = friction
        s["restitution"] = restitution
        self._shapes.append(s)
        return len(self._shapes) - 1

    # -- convenience (physxcoopa/tests/support/physics_fixtures.h:37-118) --
    def static_box(self, position, half_extents, orientation=(0, 0, 0, 1), **mat):
        i = self.add_body(PXG_STATIC, position, orientation, mass=1.0)
        self.add_shape(i, PXG_BOX, half_extents=half_extents, **mat)
        return i

    def dynamic_box(self, position, half_extents, orientation=(0, 0, 0, 1), mass=1.0, **kw):
        mat = {k: kw.pop(k) for k in list(kw) if k in ("friction", "restitution", "friction_combine",
                                                       "restitution_combine", "is_trigger", "layer")}
        i = self.add_body(PXG_DYNAMIC, position, orientation, mass=mass,
                          inv_inertia=box_inverse_inertia(half_extents, mass), **kw)
        self.add_shape(i, PXG_BOX, half_extents=half_extents, **mat)
        return i

    def dynamic_sphere(self, position, radius, mass=1.0, **kw):
        mat = {k: kw.pop(k) for k in list(kw) if k in ("friction", "restitution", "friction_combine",
                                                       "restitution_combine", "is_trigger", "layer")}
        i = self.add_body(PXG_DYNAMIC, position, mass=mass, inv_inertia=sphere_inverse_inertia(radius, mass), **kw)
        self.add_shape(i, PXG_SPHERE, radius=radius, **mat)
        return i

    def dynamic_capsule(self, position, radius, half_height, axis=2, orientation=(0, 0, 0, 1), mass=1.0, **kw):
        mat = {k: kw.pop(k) for k in list(kw) if k in ("friction", "restitution", "friction_combine",
                                                       "restitution_combine", "is_trigger", "layer")}
        raw = capsule_inverse_inertia(radius, half_height, mass)
        if axis == 0:
            local = (raw[2], raw[1], raw[0])
        elif axis == 1:
            local = (raw[0], raw[2], raw[1])
        else:
            local = raw
        i = self.add_body(PXG_DYNAMIC, position, orientation, mass=mass, inv_inertia=local, **kw)
        self.add_shape(i, PXG_CAPSULE, capsule_radius=radius, capsule_half_height=half_height, capsule_axis=axis, **mat)
        return i

    def arrays(self):
        bodies = np.array(self._bodies, dtype=BODY_DTYPE) if self._bodies else np.zeros(0, BODY_DTYPE)
        shapes = np.array(self._shapes, dtype=SHAPE_DTYPE) if self._shapes else np.zeros(0, SHAPE_DTYPE)
        return np.ascontiguousarray(bodies), np.ascontiguousarray(shapes)


class Scene:
    """A generated scene: ABI arrays plus the solver settings it is quoted with."""

    def __init__(self, name, bodies, shapes, config=None, steps=300):
        self.name = name
        self.bodies = bodies
        self.shapes = shapes
        self.config = dict(config or {})
        self.steps = steps

    @property
    def n_dynamic(self):
        return int(np.count_nonzero(self.bodies["type"] == PXG_DYNAMIC))


def config1_scene(seed=1):
    """64 dynamic unit boxes (half 0.5 m, 1 kg) on a 4x4x4 grid, 1.2 m spacing, bottom faces 1 m
    above a static 50x50x1 m ground; jitter U(+-0.05 m) in position and U(+-5 deg) in yaw.
    Friction 0.5, restitution 0, 8 velocity and 3 position iterations, 300 steps."""
    rng = np.random.default_rng(seed)
    sb = SceneBuilder()
    sb.static_box((0, 0, -0.5), (25, 25, 0.5), friction=0.5)
    for k in range(4):
        for j in range(4):
            for i in range(4):
                jit = rng.uniform(-0.05, 0.05, 3)
                yaw = math.radians(rng.uniform(-5.0, 5.0))
                pos = ((i - 1.5) * 1.2 + jit[0], (j - 1.5) * 1.2 + jit[1], 1.5 + k * 1.2 + jit[2])
                sb.dynamic_box(pos, (0.5, 0.5, 0.5), quat_axis_angle((0, 0, 1), yaw), mass=1.0, friction=0.5)
    b, s = sb.arrays()
    return Scene("config1_64_boxes", b, s, dict(velocity_iterations=8, position_iterations=3), steps=300)


def _pile_shape(rng):
    """One body of the 40% box / 40% sphere / 20% capsule mix: (kind, dims, bounding radius)."""
    u = rng.random()
    if u < 0.4:
        he = tuple(float(F(x)) for x in rng.uniform(0.2, 0.6, 3))
        return PXG_BOX, he, math.sqrt(sum(h * h for h in he))
    if u < 0.8:
        r = float(F(rng.uniform(0.2, 0.5)))
        return PXG_SPHERE, r, r
    return PXG_CAPSULE, (0.2, 0.4), 0.6


def _pile_mass(kind, dims, density=1000.0):
    rho = F(density)
    if kind == PXG_BOX:
        hx, hy, hz = (F(d) for d in dims)
        return (((F(8.0) * hx) * hy) * hz) * rho
    if kind == PXG_SPHERE:
        r = F(dims)
        return ((((F(4.0 / 3.0) * PI_F) * r) * r) * r) * rho
    r, hh = F(dims[0]), F(dims[1])
    return (((PI_F * r) * r) * (F(2.0) * hh) + (((F(4.0 / 3.0) * PI_F) * r) * r) * r) * rho


def _add_pile_body(sb, kind, dims, pos, quat, friction, restitution):
    m = _pile_mass(kind, dims)
    if kind == PXG_BOX:
        return sb.dynamic_box(pos, dims, quat, mass=m, friction=friction, restitution=restitution)
    if kind == PXG_SPHERE:
        i = sb.dynamic_sphere(pos, dims, mass=m, friction=friction, restitution=restitution)
        sb._bodies[i]["orientation"] = quat
        sb._bodies[i]["prev_orientation"] = quat
        return i
    return sb.dynamic_capsule(pos, dims[0], dims[1], 2, quat, mass=m, friction=friction, restitution=restitution)


def config2_scene(seed=2, n=16384, lattice=(64, 64, 40)):
    """The large mixed pile: n dynamic bodies (40% boxes half U[0.2,0.6]^3, 40% spheres r
    U[0.2,0.5], 20% Z capsules r 0.2 / half-height 0.4), density 1000 kg/m^3, uniform random
    orientation, one body per 1 m cell of a 64x64x40 lattice whose floor is 1 m above a static
    200x200x1 m ground. Friction U[0.3,0.8], restitution U[0,0.2], Average combine (the ground
    samples its own pair once). 8 velocity / 3 position iterations, 600 steps.

    No initial overlap: boxes up to 0.6 m half-extent and 1.2 m capsules are larger than a 1 m
    cell, so a body whose bounding sphere would overlap an already-placed neighbour's is moved
    to another free cell (seeded rejection sampling)."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = lattice
    ncell = nx * ny * nz
    assert n <= ncell
    sb = SceneBuilder()
    gf, gr = float(F(rng.uniform(0.3, 0.8))), float(F(rng.uniform(0.0, 0.2)))
    sb.static_box((0, 0, -0.5), (100, 100, 0.5), friction=gf, restitution=gr)
    occupied = {}
    perm = rng.permutation(ncell)
    cursor = 0
    for _ in range(n):
        kind, dims, brad = _pile_shape(rng)
        quat = random_quaternion(rng)
        fr, re = float(F(rng.uniform(0.3, 0.8))), float(F(rng.uniform(0.0, 0.2)))
        while True:
            c = int(perm[cursor])
            cursor += 1
            ix, iy, iz = c % nx, (c // nx) % ny, c // (nx * ny)
            center = (ix - nx / 2 + 0.5, iy - ny / 2 + 0.5, 1.0 + iz + 0.5)
            ok = True
            for dx in (-2, -1, 0, 1, 2):
                for dy in (-2, -1, 0, 1, 2):
                    for dz in (-2, -1, 0, 1, 2):
                        o = occupied.get((ix + dx, iy + dy, iz + dz))
                        if o is None:
                            continue
                        d2 = sum((center[t] - o[0][t]) ** 2 for t in range(3))
                        if d2 < (brad + o[1]) ** 2:
                            ok = False
            if ok:
                break
        occupied[(ix, iy, iz)] = (center, brad)
        _add_pile_body(sb, kind, dims, center, quat, fr, re)
    b, s = sb.arrays()
    return Scene("config2_pile_%d" % n, b, s, dict(velocity_iterations=8, position_iterations=3), steps=600)


def pyramid_scene(n_bodies, friction=0.5):
    """A 2-D box pyramid of n_bodies unit boxes (base width ceil(sqrt(2N))) on a 20x20x1 m ground."""
    sb = SceneBuilder()
    sb.static_box((0, 0, -0.5), (10, 10, 0.5), friction=friction)
    base = int(math.ceil(math.sqrt(2 * n_bodies)))
    placed = 0
    row = 0
    while placed < n_bodies and base - row > 0:
        width = base - row
        for i in range(width):
            if placed >= n_bodies:
                break
            x = (i - (width - 1) / 2.0) * 1.0
            sb.dynamic_box((x, 0.0, 0.5 + row * 1.0), (0.5, 0.5, 0.5), mass=1.0, friction=friction)
            placed += 1
        row += 1
    return sb.arrays()


def drop_scene(n_bodies, rng):
    """n_bodies of the pile shape mix dropped in a 10x10x20 m column over a 20x20x1 m ground."""
    sb = SceneBuilder()
    sb.static_box((0, 0, -0.5), (10, 10, 0.5), friction=0.5)
    # 1.25 m cells in an 8x8 column: no initial overlap for the largest body (bounding radius 1.04)
    cells = rng.permutation(8 * 8 * 16)[:n_bodies]
    for c in cells:
        ix, iy, iz = c % 8, (c // 8) % 8, c // 64
        kind, dims, _ = _pile_shape(rng)
        quat = random_quaternion(rng)
        pos = ((ix - 3.5) * 1.25, (iy - 3.5) * 1.25, 1.2 + iz * 1.25)
        _add_pile_body(sb, kind, dims, pos, quat, 0.5, 0.0)
    return sb.arrays()


def hull_pile_scene(hulls, n=40, seed=4, spread=1.5, others=True):
    """Convex-hull bodies (cycling through `hulls`, a list of _abi.Hull; 2 kg, the inertia of their
    bounding box) dropped with random orientations over a 20x20x1 m ground box, with spheres,
    boxes and capsules among them when `others`."""
    rng = np.random.default_rng(seed)
    sb = SceneBuilder()
    sb.static_box((0, 0, -0.5), (10, 10, 0.5), friction=0.6)
    for i in range(n):
        pos = (rng.uniform(-spread, spread), rng.uniform(-spread, spread), 0.7 + 0.9 * i)
        q = random_quaternion(rng)
        kind = i % 5 if others else 0
        if kind in (0, 1, 2):
            k = i % len(hulls)
            h = hulls[k]
            ext = (np.asarray(h.bounds6[3:], np.float64) - np.asarray(h.bounds6[:3], np.float64)) * 0.5
            b = sb.add_body(PXG_DYNAMIC, position=pos, orientation=q, mass=2.0,
                            inv_inertia=box_inverse_inertia(ext, 2.0))
            sb.add_shape(b, PXG_CONVEX_HULL, radius=1.0, capsule_axis=k, friction=0.6)
        elif kind == 3:
            sb.dynamic_sphere(pos, rng.uniform(0.2, 0.4), mass=1.5, friction=0.5)
        else:
            sb.dynamic_capsule(pos, 0.15, 0.3, axis=int(rng.integers(0, 3)), orientation=q, mass=1.5, friction=0.5)
    return sb.arrays()


# ---- triangle meshes (the meshes of physxcoopa/tests/narrowphase_test.cpp:200-490) ----

def tiled_floor(n):
    """n x n quads over [-2, 2]^2 at z = 0, two triangles each (narrowphase_test.cpp:314-327)."""
    step = np.float32(4.0) / np.float32(n)
    v = [(np.float32(-2.0) + np.float32(i) * step, np.float32(-2.0) + np.float32(j) * step, 0.0)
         for j in range(n + 1) for i in range(n + 1)]
    idx = []
    for j in range(n):
        for i in range(n):
            a = j * (n + 1) + i
            idx += [a, a + 1, a + n + 2, a, a + n + 2, a + n + 1]
    return TriangleMesh(v, idx)


def two_triangle_floor():
    """A 4 x 4 m floor of two triangles sharing the diagonal x == y."""
    return TriangleMesh([(-2, -2, 0), (2, -2, 0), (2, 2, 0), (-2, 2, 0)], [0, 1, 2, 0, 2, 3])


def ledge():
    """A flat top (x in [-2, 0], z = 0, facing +z) and the wall under its edge (x = 0, facing +x),
    wall triangles first (narrowphase_test.cpp:307-312)."""
    return TriangleMesh([(-2, -2, 0), (0, -2, 0), (0, 2, 0), (-2, 2, 0), (0, -2, -2), (0, 2, -2)],
                        [1, 4, 5, 1, 5, 2, 0, 1, 2, 0, 2, 3])


def bumpy_terrain(cells=16, cell=0.5, amplitude=0.4):
    """A cells x cells heightfield of z = amplitude sin(1.3 x) cos(0.9 y) in float32
    (narrowphase_test.cpp:200-216), as a triangle soup."""
    v = []
    for j in range(cells + 1):
        for i in range(cells + 1):
            x, y = np.float32(i * cell), np.float32(j * cell)
            v.append((x, y, np.float32(amplitude) * np.float32(math.sin(x * 1.3)) * np.float32(math.cos(y * 0.9))))
    idx = []
    for j in range(cells):
        for i in range(cells):
            a = j * (cells + 1) + i
            b, c = a + 1, a + cells + 1
            idx += [a, b, c + 1, a, c + 1, c]
    return TriangleMesh(v, idx)


def mesh_shape_on(sb, mesh_index, position=(0, 0, 0), orientation=(0, 0, 0, 1), scale=1.0, btype=PXG_STATIC,
                  friction=0.6, local_center=(0, 0, 0), local_rotation=(0, 0, 0, 1)):
    """A static (or kinematic) body carrying a PXG_TRIANGLE_MESH shape of table entry mesh_index."""
    b = sb.add_body(btype, position, orientation, mass=1.0)
    sb.add_shape(b, PXG_TRIANGLE_MESH, radius=scale, capsule_axis=mesh_index, friction=friction,
                 local_center=local_center, local_rotation=local_rotation)
    return b


def terrain_pile_scene(hulls=None, n=48, seed=11, mesh_position=(-4.0, -4.0, 0.0), mesh_orientation=(0, 0, 0, 1)):
    """Spheres, boxes, capsules and (given `hulls`) convex hulls dropped with random orientations
    onto the bumpy terrain mesh (table entry 0; 8 x 8 m placed at mesh_position), for lockstep
    tests of the mesh narrowphase."""
    rng = np.random.default_rng(seed)
    sb = SceneBuilder()
    mesh_shape_on(sb, 0, position=mesh_position, orientation=mesh_orientation)
    kinds = 4 if hulls else 3
    for i in range(n):
        pos = (rng.uniform(-2.5, 2.5), rng.uniform(-2.5, 2.5), 1.0 + 0.35 * i)
        q = random_quaternion(rng)
        kind = i % kinds
        if kind == 0:
            sb.dynamic_sphere(pos, rng.uniform(0.15, 0.35), mass=1.0, friction=0.5)
        elif kind == 1:
            he = tuple(rng.uniform(0.12, 0.35, 3))
            sb.dynamic_box(pos, he, orientation=q, mass=1.5, friction=0.6)
        elif kind == 2:
            sb.dynamic_capsule(pos, rng.uniform(0.1, 0.2), rng.uniform(0.15, 0.35), axis=int(rng.integers(0, 3)),
                               orientation=q, mass=1.2, friction=0.5)
        else:
            k = i % len(hulls)
            h = hulls[k]
            ext = (np.asarray(h.bounds6[3:], np.float64) - np.asarray(h.bounds6[:3], np.float64)) * 0.5
            b = sb.add_body(PXG_DYNAMIC, position=pos, orientation=q, mass=2.0, inv_inertia=box_inverse_inertia(ext, 2.0))
            sb.add_shape(b, PXG_CONVEX_HULL, radius=1.0, capsule_axis=k, friction=0.6)
    return sb.arrays()


def config3_scenes(seed=3, n_scenes=4096, n_min=8, n_max=256):
    """Independent scenes with N ~ U{8..256} dynamic bodies: even scenes are box pyramids, odd
    scenes random drops of the pile shape mix. 8 velocity / 3 position iterations, 300 steps.
    Returns (bodies, shapes, body_offsets, shape_offsets, settings)."""
    rng = np.random.default_rng(seed)
    bl, sl = [], []
    for k in range(n_scenes):
        nb = int(rng.integers(n_min, n_max + 1))
        b, s = pyramid_scene(nb) if k % 2 == 0 else drop_scene(nb, rng)
        bl.append(b)
        sl.append(s)
    body_off = np.zeros(n_scenes + 1, np.uint32)
    shape_off = np.zeros(n_scenes + 1, np.uint32)
    body_off[1:] = np.cumsum([len(b) for b in bl])
    shape_off[1:] = np.cumsum([len(s) for s in sl])
    bodies = np.ascontiguousarray(np.concatenate(bl))
    shapes = np.ascontiguousarray(np.concatenate(sl))
    return bodies, shapes, body_off, shape_off, dict(velocity_iterations=8, position_iterations=3)


# ---- small fixtures (physxcoopa/tests/support/physics_fixtures.h) ----

def ground(sb, material=None):
    """Static 8 x 8 m slab whose top face is at z = 0 (physics_fixtures.h:37-44)."""
    return sb.static_box((0, 0, -0.1), (4, 4, 0.1), **(material or {}))


def sphere_rain_scene(n=80, seed=1234, radius=0.4):
    """80 unit spheres dropped over the fixture ground (physics_world_test.cpp:177-207 layout)."""
    rng = np.random.default_rng(seed)
    sb = SceneBuilder()
    ground(sb)
    for _ in range(n):
        p = (rng.uniform(-3, 3), rng.uniform(-3, 3), rng.uniform(2, 8))
        sb.dynamic_sphere(p, radius, mass=1.0)
    return sb.arrays()


def mixed_pile_scene(n=200, seed=7, spread=4.0):
    """A small mixed pile (boxes, spheres, capsules, random orientations) over the fixture ground."""
    rng = np.random.default_rng(seed)
    sb = SceneBuilder()
    sb.static_box((0, 0, -0.5), (10, 10, 0.5), friction=0.6)
    side = int(math.ceil(math.sqrt(n / 4.0)))
    for k in range(n):
        ix, iy, iz = k % side, (k // side) % side, k // (side * side)
        kind, dims, _ = _pile_shape(rng)
        quat = random_quaternion(rng)
        pos = ((ix - side / 2) * 1.3 * spread / 4.0, (iy - side / 2) * 1.3 * spread / 4.0, 1.2 + iz * 1.3)
        fr, re = float(F(rng.uniform(0.3, 0.8))), float(F(rng.uniform(0.0, 0.2)))
        _add_pile_body(sb, kind, dims, pos, quat, fr, re)
    return sb.arrays()


# ---- ragdolls (joints) ----
# A 15-bone capsule ragdoll standing along +Z, pelvis centre at the origin of its own frame.
# (name, capsule axis, radius, half height, centre, mass)
RAGDOLL_BONES = [
    ("pelvis", 0, 0.12, 0.08, (0.0, 0.0, 0.0), 9.0),
    ("abdomen", 2, 0.11, 0.04, (0.0, 0.0, 0.22), 8.0),
    ("chest", 0, 0.13, 0.08, (0.0, 0.0, 0.46), 12.0),
    ("neck", 2, 0.08, 0.02, (0.0, 0.0, 0.66), 1.5),
    ("head", 2, 0.10, 0.04, (0.0, 0.0, 0.86), 6.5),
    ("upper_arm_l", 0, 0.08, 0.11, (0.42, 0.0, 0.46), 2.5),
    ("forearm_l", 0, 0.08, 0.10, (0.76, 0.0, 0.46), 1.8),
    ("upper_arm_r", 0, 0.08, 0.11, (-0.42, 0.0, 0.46), 2.5),
    ("forearm_r", 0, 0.08, 0.10, (-0.76, 0.0, 0.46), 1.8),
    ("thigh_l", 2, 0.10, 0.14, (0.14, 0.0, -0.36), 7.0),
    ("shin_l", 2, 0.08, 0.15, (0.14, 0.0, -0.82), 4.0),
    ("foot_l", 1, 0.08, 0.04, (0.14, 0.06, -1.10), 1.2),
    ("thigh_r", 2, 0.10, 0.14, (-0.14, 0.0, -0.36), 7.0),
    ("shin_r", 2, 0.08, 0.15, (-0.14, 0.0, -0.82), 4.0),
    ("foot_r", 1, 0.08, 0.04, (-0.14, 0.06, -1.10), 1.2),
]
# (kind, parent, child, junction point, axis (hinge axis / twist axis), limits)
#   cone: (swing, twist_min, twist_max) in degrees; hinge: (min, max) in degrees
RAGDOLL_JOINTS = [
    ("cone", "pelvis", "abdomen", (0.0, 0.0, 0.11), (0, 0, 1), (30.0, -20.0, 20.0)),
    ("cone", "abdomen", "chest", (0.0, 0.0, 0.33), (0, 0, 1), (25.0, -15.0, 15.0)),
    ("cone", "chest", "neck", (0.0, 0.0, 0.58), (0, 0, 1), (40.0, -40.0, 40.0)),
    ("hinge", "neck", "head", (0.0, 0.0, 0.74), (1, 0, 0), (-30.0, 40.0)),
    ("cone", "chest", "upper_arm_l", (0.25, 0.0, 0.46), (1, 0, 0), (70.0, -45.0, 45.0)),
    ("hinge", "upper_arm_l", "forearm_l", (0.59, 0.0, 0.46), (0, 0, 1), (0.0, 140.0)),
    ("cone", "chest", "upper_arm_r", (-0.25, 0.0, 0.46), (-1, 0, 0), (70.0, -45.0, 45.0)),
    ("hinge", "upper_arm_r", "forearm_r", (-0.59, 0.0, 0.46), (0, 0, 1), (-140.0, 0.0)),
    ("cone", "pelvis", "thigh_l", (0.14, 0.0, -0.14), (0, 0, -1), (60.0, -30.0, 30.0)),
    ("hinge", "thigh_l", "shin_l", (0.14, 0.0, -0.59), (1, 0, 0), (-140.0, 0.0)),
    ("hinge", "shin_l", "foot_l", (0.14, 0.0, -1.03), (1, 0, 0), (-30.0, 30.0)),
    ("cone", "pelvis", "thigh_r", (-0.14, 0.0, -0.14), (0, 0, -1), (60.0, -30.0, 30.0)),
    ("hinge", "thigh_r", "shin_r", (-0.14, 0.0, -0.59), (1, 0, 0), (-140.0, 0.0)),
    ("hinge", "shin_r", "foot_r", (-0.14, 0.0, -1.03), (1, 0, 0), (-30.0, 30.0)),
]


def _seg_dist(p1, q1, p2, q2):
    """Distance between segments p1q1 and p2q2 (float64, for the rest-pose overlap test)."""
    d1, d2, r = q1 - p1, q2 - p2, p1 - p2
    a, e, f = d1 @ d1, d2 @ d2, d2 @ r
    c = d1 @ r
    b = d1 @ d2
    den = a * e - b * b
    s = np.clip((b * f - c * e) / den, 0.0, 1.0) if den > 1e-12 else 0.0
    t = (b * s + f) / e if e > 1e-12 else 0.0
    if t < 0.0:
        t, s = 0.0, np.clip(-c / a, 0.0, 1.0)
    elif t > 1.0:
        t, s = 1.0, np.clip((b - c) / a, 0.0, 1.0)
    return float(np.linalg.norm((p1 + d1 * s) - (p2 + d2 * t)))


def add_ragdoll(sb, origin, orientation, rng=None, friction=0.6):
    """Adds a 15-bone ragdoll (70 kg) at `origin` rotated by `orientation` (x, y, z, w). Hinge limits
    are jittered by up to 10 degrees with `rng`. Returns (joint records, ignored body pairs): every
    joint has collide_connected off, and bones that overlap at rest without sharing a joint are
    returned as pairs for set_pair_collision(a, b, false) (world.h:527-532)."""
    from . import joints as J
    q = np.array(orientation, np.float64)
    idx = {}
    segs = {}
    for name, axis, r, hh, c, m in RAGDOLL_BONES:
        c = np.array(c, np.float64)
        pos = np.asarray(origin, np.float64) + J.quat_rotate(q, c).astype(np.float64)
        idx[name] = sb.dynamic_capsule(tuple(float(F(x)) for x in pos), r, hh, axis=axis,
                                       orientation=tuple(float(F(x)) for x in q), mass=m, friction=friction)
        e = np.zeros(3)
        e[axis] = hh
        segs[name] = (c - e, c + e, r)
    b, _ = sb.arrays()
    recs = []
    centre = {name: np.array(c, np.float64) for name, _, _, _, c, _ in RAGDOLL_BONES}
    jitter = (lambda: float(rng.uniform(-10.0, 10.0))) if rng is not None else (lambda: 0.0)
    for kind, pa, ch, jp, ax, lim in RAGDOLL_JOINTS:
        jp = np.array(jp, np.float64)
        la = tuple(float(F(x)) for x in jp - centre[pa])
        lb = tuple(float(F(x)) for x in jp - centre[ch])
        if kind == "cone":
            sw, tmin, tmax = lim
            recs.append(J.cone_twist(b, idx[pa], idx[ch], la, lb, ax, F(math.radians(sw)), F(math.radians(tmin)),
                                     F(math.radians(tmax))))
        else:
            lo, hi = lim
            j = jitter()
            recs.append(J.hinge(b, idx[pa], idx[ch], la, lb, ax, ax, True, F(math.radians(lo + min(j, 0.0))),
                                F(math.radians(hi + max(j, 0.0)))))
    jointed = {(min(idx[p], idx[c]), max(idx[p], idx[c])) for _, p, c, _, _, _ in RAGDOLL_JOINTS}
    ignored = []
    names = [n for n, *_ in RAGDOLL_BONES]
    for i, na in enumerate(names):
        for nb in names[i + 1:]:
            key = (min(idx[na], idx[nb]), max(idx[na], idx[nb]))
            if key in jointed:
                continue
            (p1, q1, r1), (p2, q2, r2) = segs[na], segs[nb]
            if _seg_dist(p1, q1, p2, q2) < r1 + r2 + 0.05:
                ignored.append(key)
    return recs, ignored


def ragdoll_scene(seed=5, n_ragdolls=1, n_boxes=32, spread=3.0):
    """The ragdoll scene: `n_ragdolls` 15-bone capsule ragdolls (70 kg each; cone-twist at the
    spine, neck, shoulders and hips, hinges at the head, elbows, knees and ankles with jittered
    limits) dropped from 2 m with random orientations onto `n_boxes` random boxes (half extents
    U[0.1, 0.4] m) over a 10 x 10 x 1 m ground box. dt 1/120, 12 velocity iterations.
    Returns (Scene, joint records, ignored pairs)."""
    rng = np.random.default_rng(seed)
    sb = SceneBuilder()
    sb.static_box((0, 0, -0.5), (5, 5, 0.5), friction=0.6)
    for _ in range(n_boxes):
        he = tuple(float(F(x)) for x in rng.uniform(0.1, 0.4, 3))
        pos = (float(F(rng.uniform(-4, 4))), float(F(rng.uniform(-4, 4))), float(F(he[2] + 0.001)))
        yaw = quat_axis_angle((0.0, 0.0, 1.0), float(rng.uniform(0, 2 * math.pi)))
        mass = float(F(8.0 * he[0] * he[1] * he[2] * 500.0))
        sb.dynamic_box(pos, he, orientation=tuple(float(F(x)) for x in yaw), mass=mass, friction=0.6)
    recs, ignored = [], []
    for k in range(n_ragdolls):
        ang = 2 * math.pi * k / max(1, n_ragdolls)
        rad = 0.0 if n_ragdolls == 1 else spread
        origin = (rad * math.cos(ang), rad * math.sin(ang), 2.0 + 1.2)
        r, ig = add_ragdoll(sb, origin, random_quaternion(rng), rng)
        recs.extend(r)
        ignored.extend(ig)
    b, s = sb.arrays()
    cfg = dict(fixed_dt=float(F(1.0) / F(120.0)), velocity_iterations=12)
    return Scene("ragdolls_%d" % n_ragdolls, b, s, cfg, steps=600), recs, ignored


def config5_scenes(seed=5, n_scenes=2048, n_boxes=32):
    """The batched ragdoll workload: `n_scenes` independent ragdoll scenes (one ragdoll, 32 boxes,
    the 10 x 10 x 1 m ground; see ragdoll_scene). Returns (bodies, shapes, body_offsets,
    shape_offsets, joints, joint_offsets, ignored pairs, pair_offsets, settings); joints and pairs
    use scene-local body indices."""
    from . import joints as J
    bl, sl, jl, il = [], [], [], []
    for k in range(n_scenes):
        sc, recs, ig = ragdoll_scene(seed=seed * 100003 + k, n_ragdolls=1, n_boxes=n_boxes)
        bl.append(sc.bodies)
        sl.append(sc.shapes)
        jl.append(J.table(recs))
        il.append(np.asarray(ig, np.uint32).reshape(-1, 2))
    off = lambda xs: np.concatenate([[0], np.cumsum([len(x) for x in xs])]).astype(np.uint32)
    return (np.ascontiguousarray(np.concatenate(bl)), np.ascontiguousarray(np.concatenate(sl)), off(bl), off(sl),
            np.ascontiguousarray(np.concatenate(jl)), off(jl), np.ascontiguousarray(np.concatenate(il)), off(il),
            dict(fixed_dt=float(F(1.0) / F(120.0)), velocity_iterations=12))
