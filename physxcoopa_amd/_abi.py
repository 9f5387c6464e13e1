"""Data layouts of the pxg C ABI (include/pxg.h) as numpy structured dtypes and ctypes structs.

The same layouts are used to feed the GPU library and the CPU oracle, so one scene buffer can
drive both. Field order and sizes must match include/pxg.h exactly; tests/test_abi.py checks the
sizes against the compiled library.
"""
import ctypes

import numpy as np

PXG_OK = 0
PXG_ERR_INVALID, PXG_ERR_CAPACITY, PXG_ERR_HIP, PXG_ERR_STATE, PXG_ERR_DEVICE = 1, 2, 3, 4, 5
PXG_STATIC, PXG_KINEMATIC, PXG_DYNAMIC = 0, 1, 2
PXG_INTERPOLATE, PXG_INTERPOLATE_NONE = 0, 1
PXG_SPHERE, PXG_BOX, PXG_CAPSULE, PXG_CONVEX_HULL, PXG_TRIANGLE_MESH = 0, 1, 2, 3, 4
PXG_AVERAGE, PXG_MINIMUM, PXG_MAXIMUM, PXG_MULTIPLY = 0, 1, 2, 3
PXG_PHASE_FORCES_COLLIDE, PXG_PHASE_SOLVE_INTEGRATE = 0, 1
PXG_ORDER_REFERENCE, PXG_ORDER_COLOURED, PXG_ORDER_TILED = 0, 1, 2

BODY_DTYPE = np.dtype([
    ("type", "<u4"), ("alive", "<u4"), ("awake", "<u4"), ("use_gravity", "<u4"),
    ("constraints", "<u4"), ("generation", "<u4"), ("interpolation", "<u4"), ("reserved1", "<u4"),
    ("position", "<f4", 3), ("sleep_timer", "<f4"),
    ("orientation", "<f4", 4),
    ("linear_velocity", "<f4", 3), ("mass", "<f4"),
    ("angular_velocity", "<f4", 3), ("inv_mass", "<f4"),
    ("force", "<f4", 3), ("linear_drag", "<f4"),
    ("torque", "<f4", 3), ("angular_drag", "<f4"),
    ("inv_inertia_local", "<f4", 3), ("reserved2", "<f4"),
    ("inv_inertia_world", "<f4", 9),
    ("prev_position", "<f4", 3), ("reserved3", "<f4"),
    ("prev_orientation", "<f4", 4),
])
assert BODY_DTYPE.itemsize == 212

POSE_DTYPE = np.dtype([
    ("position", "<f4", 3), ("type", "<u4"), ("orientation", "<f4", 4),
    ("prev_position", "<f4", 3), ("interpolation", "<u4"), ("prev_orientation", "<f4", 4),
])
assert POSE_DTYPE.itemsize == 64

SHAPE_DTYPE = np.dtype([
    ("owner", "<u4"), ("type", "<u4"), ("enabled", "<u4"), ("is_trigger", "<u4"),
    ("layer", "<u4"), ("capsule_axis", "<i4"), ("friction_combine", "<u4"), ("restitution_combine", "<u4"),
    ("local_center", "<f4", 3), ("radius", "<f4"),
    ("local_rotation", "<f4", 4),
    ("half_extents", "<f4", 3), ("capsule_radius", "<f4"),
    ("capsule_half_height", "<f4"), ("dynamic_friction", "<f4"), ("static_friction", "<f4"), ("restitution", "<f4"),
])
assert SHAPE_DTYPE.itemsize == 96

CONTACT_DTYPE = np.dtype([
    ("position", "<f4", 3), ("penetration", "<f4"), ("separation", "<f4"), ("normal_impulse", "<f4"),
    ("tangent_impulse", "<f4", 2), ("feature_id", "<u4"), ("restitution_bias", "<f4"), ("reserved", "<f4", 2),
])
assert CONTACT_DTYPE.itemsize == 48

MANIFOLD_DTYPE = np.dtype([
    ("a", "<u4"), ("b", "<u4"), ("shape_a", "<u4"), ("shape_b", "<u4"),
    ("normal", "<f4", 3), ("friction", "<f4"),
    ("restitution", "<f4"), ("count", "<u4"), ("is_trigger", "<u4"), ("valid", "<u4"),
    ("points", CONTACT_DTYPE, 4),
])
assert MANIFOLD_DTYPE.itemsize == 240

PXG_EVENT_ENTER, PXG_EVENT_STAY, PXG_EVENT_EXIT = 0, 1, 2

EVENT_DTYPE = np.dtype([("a", "<u4"), ("b", "<u4"), ("type", "<u4"), ("is_trigger", "<u4")])
assert EVENT_DTYPE.itemsize == 16

RAY_DTYPE = np.dtype([("origin", "<f4", 3), ("max_distance", "<f4"), ("direction", "<f4", 3), ("reserved", "<f4")])
assert RAY_DTYPE.itemsize == 32
PXG_NO_BODY = 0xFFFFFFFF
FILTER_DTYPE = np.dtype([("layer_mask", "<u4"), ("include_triggers", "<u4"), ("ignore_body", "<u4"), ("reserved", "<u4")])
assert FILTER_DTYPE.itemsize == 16
HIT_DTYPE = np.dtype([("point", "<f4", 3), ("distance", "<f4"), ("normal", "<f4", 3), ("hit", "<u4"), ("body", "<u4"),
                      ("shape_index", "<u4"), ("started_inside", "<u4"), ("reserved", "<u4")])
assert HIT_DTYPE.itemsize == 48


def rays(origins, directions, max_distance=np.finfo(np.float32).max):
    """A RAY_DTYPE array from (n, 3) origins and directions (and one or n max distances)."""
    o = np.asarray(origins, np.float32).reshape(-1, 3)
    r = np.zeros(len(o), RAY_DTYPE)
    r["origin"] = o
    r["direction"] = np.asarray(directions, np.float32).reshape(-1, 3)
    r["max_distance"] = max_distance
    return r


OVERLAP_DTYPE = np.dtype([("type", "<u4"), ("radius", "<f4"), ("center", "<f4", 3), ("half_extents", "<f4", 3),
                          ("orientation", "<f4", 4), ("a", "<f4", 3), ("b", "<f4", 3)])
assert OVERLAP_DTYPE.itemsize == 72


def overlap_sphere(center, radius):
    q = np.zeros(1, OVERLAP_DTYPE)
    q["type"], q["center"], q["radius"], q["orientation"] = PXG_SPHERE, center, radius, (0, 0, 0, 1)
    return q


def overlap_box(center, half_extents, orientation=(0, 0, 0, 1)):
    q = np.zeros(1, OVERLAP_DTYPE)
    q["type"], q["center"], q["half_extents"], q["orientation"] = PXG_BOX, center, half_extents, orientation
    return q


def overlap_capsule(a, b, radius):
    q = np.zeros(1, OVERLAP_DTYPE)
    q["type"], q["a"], q["b"], q["radius"], q["orientation"] = PXG_CAPSULE, a, b, radius, (0, 0, 0, 1)
    return q


CAST_DTYPE = np.dtype([("shape", SHAPE_DTYPE), ("position", "<f4", 3), ("max_distance", "<f4"), ("orientation", "<f4", 4),
                       ("direction", "<f4", 3), ("reserved", "<f4")])
assert CAST_DTYPE.itemsize == 144

PEN_DTYPE = np.dtype([("normal", "<f4", 3), ("depth", "<f4"), ("point", "<f4", 3), ("body", "<u4"), ("shape_index", "<u4"),
                      ("reserved", "<u4")])
assert PEN_DTYPE.itemsize == 40


def probe_parts(probe):
    """(shape, position, orientation) arrays of a CAST_DTYPE record, as compute_penetration takes
    the posed query shape (its direction and max_distance are not used)."""
    c = np.ascontiguousarray(probe).reshape(1)
    return (np.ascontiguousarray(c["shape"]), np.ascontiguousarray(c["position"][0], np.float32),
            np.ascontiguousarray(c["orientation"][0], np.float32))


def sphere_cast(origin, radius, direction, max_distance):
    """sphere_cast() (world.h:911-920) as a CAST_DTYPE record."""
    c = np.zeros(1, CAST_DTYPE)
    c["shape"]["type"], c["shape"]["enabled"], c["shape"]["radius"] = PXG_SPHERE, 1, radius
    c["shape"]["local_rotation"] = (0, 0, 0, 1)
    c["position"], c["orientation"], c["direction"], c["max_distance"] = origin, (0, 0, 0, 1), direction, max_distance
    return c


def box_cast(origin, half_extents, orientation, direction, max_distance):
    """box_cast() (world.h:924-933) as a CAST_DTYPE record."""
    c = np.zeros(1, CAST_DTYPE)
    c["shape"]["type"], c["shape"]["enabled"], c["shape"]["half_extents"] = PXG_BOX, 1, half_extents
    c["shape"]["local_rotation"] = (0, 0, 0, 1)
    c["position"], c["orientation"], c["direction"], c["max_distance"] = origin, orientation, direction, max_distance
    return c


def capsule_cast(origin, radius, half_height, axis, orientation, direction, max_distance):
    """capsule_cast() (world.h:938-950) as a CAST_DTYPE record."""
    c = np.zeros(1, CAST_DTYPE)
    sh = c["shape"]
    sh["type"], sh["enabled"], sh["capsule_radius"], sh["capsule_half_height"] = PXG_CAPSULE, 1, radius, half_height
    sh["capsule_axis"], sh["local_rotation"] = axis, (0, 0, 0, 1)
    c["position"], c["orientation"], c["direction"], c["max_distance"] = origin, orientation, direction, max_distance
    return c


def query_filter(layer_mask=0xFFFFFFFF, include_triggers=True, ignore_body=PXG_NO_BODY):
    f = np.zeros(1, FILTER_DTYPE)
    f["layer_mask"] = layer_mask
    f["include_triggers"] = 1 if include_triggers else 0
    f["ignore_body"] = ignore_body
    return f


PXG_HINGE, PXG_BALL, PXG_CONE_TWIST = 0, 1, 2

JOINT_DTYPE = np.dtype([
    ("type", "<u4"), ("body_a", "<u4"), ("body_b", "<u4"), ("valid", "<u4"),
    ("enabled", "<u4"), ("use_limits", "<u4"), ("collide_connected", "<u4"), ("reserved0", "<u4"),
    ("local_anchor_a", "<f4", 3), ("min_angle", "<f4"),
    ("local_anchor_b", "<f4", 3), ("max_angle", "<f4"),
    ("local_axis_a", "<f4", 3), ("swing_limit", "<f4"),
    ("local_axis_b", "<f4", 3), ("twist_min", "<f4"),
    ("rest_relative_rotation", "<f4", 4),
    ("twist_max", "<f4"), ("limit_impulse", "<f4"), ("swing_impulse", "<f4"), ("reserved1", "<f4"),
    ("point_impulse", "<f4", 3), ("reserved2", "<f4"),
    ("axis_impulse", "<f4", 2), ("reserved3", "<f4", 2),
])
assert JOINT_DTYPE.itemsize == 160


BATCH_STATS_DTYPE = np.dtype([
    ("pairs", "<u4"), ("manifolds", "<u4"), ("points", "<u4"), ("colours", "<u4"),
    ("active_manifolds", "<u4"), ("velocity_passes", "<u4"), ("error", "<u4"), ("ordered", "<u4"),
    ("residual", "<f4"), ("awake_bodies", "<u4"), ("learned_predictions", "<u4"), ("velocity_passes_total", "<u4"),
])
assert BATCH_STATS_DTYPE.itemsize == 48


class Config(ctypes.Structure):
    """pxg_config (include/pxg.h)."""
    _fields_ = [
        ("velocity_iterations", ctypes.c_uint32), ("relax_iterations", ctypes.c_uint32),
        ("position_iterations", ctypes.c_uint32), ("position_correction", ctypes.c_float),
        ("linear_slop", ctypes.c_float), ("max_linear_correction", ctypes.c_float),
        ("restitution_threshold", ctypes.c_float), ("sleep_linear", ctypes.c_float),
        ("sleep_angular", ctypes.c_float), ("sleep_time", ctypes.c_float), ("aabb_margin", ctypes.c_float),
        ("max_substeps", ctypes.c_uint32), ("fixed_dt", ctypes.c_float), ("max_linear_velocity", ctypes.c_float),
        ("use_exact_quaternion_integration", ctypes.c_uint32), ("gravity", ctypes.c_float * 3),
        ("residual_mode", ctypes.c_uint32), ("residual_tolerance", ctypes.c_float),
        ("residual_max_iterations", ctypes.c_uint32), ("large_extent", ctypes.c_float),
        ("max_bodies", ctypes.c_uint32), ("max_shapes", ctypes.c_uint32), ("max_pairs", ctypes.c_uint32),
        ("max_manifolds", ctypes.c_uint32), ("learned_warm_start", ctypes.c_uint32),
        ("solve_order", ctypes.c_uint32), ("residual_check_interval", ctypes.c_uint32),
        ("cloth_substeps", ctypes.c_uint32), ("cloth_iterations", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32 * 3),
    ]


class Stats(ctypes.Structure):
    """pxg_stats (include/pxg.h)."""
    _fields_ = [
        ("pairs", ctypes.c_uint32), ("manifolds", ctypes.c_uint32), ("points", ctypes.c_uint32),
        ("colours", ctypes.c_uint32), ("active_manifolds", ctypes.c_uint32), ("velocity_passes", ctypes.c_uint32),
        ("awake_bodies", ctypes.c_uint32), ("learned_predictions", ctypes.c_uint32),
        ("residual", ctypes.c_float), ("tiled_workgroups", ctypes.c_uint32), ("steps", ctypes.c_uint64),
    ]


class Profile(ctypes.Structure):
    """pxg_profile (include/pxg.h)."""
    _fields_ = [("velocity_ms", ctypes.c_double), ("velocity_bytes", ctypes.c_double),
                ("velocity_launches", ctypes.c_uint64), ("tiled_colour_ticks", ctypes.c_uint64),
                ("tiled_stage_ticks", ctypes.c_uint64), ("tiled_barrier_ticks", ctypes.c_uint64),
                ("tiled_residual_ticks", ctypes.c_uint64), ("reserved", ctypes.c_uint64 * 1)]


def default_config(**overrides):
    """util::PhysicsConfig defaults (physxcoopa/util/config.h:22-145) plus the GPU extras."""
    c = Config()
    c.velocity_iterations = 16
    c.relax_iterations = 2
    c.position_iterations = 4
    c.position_correction = 0.2
    c.linear_slop = 0.005
    c.max_linear_correction = 0.2
    c.restitution_threshold = 0.5
    c.sleep_linear = 0.01
    c.sleep_angular = 0.02
    c.sleep_time = 0.5
    c.aabb_margin = 0.1
    c.max_substeps = 8
    c.fixed_dt = np.float32(1.0) / np.float32(60.0)
    c.max_linear_velocity = 200.0
    c.use_exact_quaternion_integration = 0
    c.gravity[0], c.gravity[1], c.gravity[2] = 0.0, 0.0, -9.81
    c.residual_mode = 0
    c.residual_tolerance = 1e-4
    c.residual_max_iterations = 256
    c.residual_check_interval = 1
    c.large_extent = 16.0
    c.learned_warm_start = 0
    c.solve_order = PXG_ORDER_REFERENCE
    c.cloth_substeps = 4
    c.cloth_iterations = 1
    for k, v in overrides.items():
        if k == "gravity":
            for i in range(3):
                c.gravity[i] = v[i]
        else:
            setattr(c, k, v)
    return c


def ptr(a):
    """ctypes void pointer to a contiguous numpy array."""
    assert a.flags["C_CONTIGUOUS"]
    return ctypes.c_void_p(a.ctypes.data)


class HullDesc(ctypes.Structure):
    """pxg_hull (include/pxg.h): a built convex hull's arrays (pointers) and sizes."""
    _fields_ = [("vertices", ctypes.c_void_p), ("n_vertices", ctypes.c_uint32), ("n_faces", ctypes.c_uint32),
                ("face_planes", ctypes.c_void_p), ("face_starts", ctypes.c_void_p), ("face_vertices", ctypes.c_void_p),
                ("edges", ctypes.c_void_p), ("n_edges", ctypes.c_uint32), ("local_min", ctypes.c_float * 3),
                ("local_max", ctypes.c_float * 3), ("centroid", ctypes.c_float * 3), ("volume", ctypes.c_float)]


class Hull:
    """A built convex hull (geometry::ConvexHull's arrays): vertices (n, 3), face planes (f, 4:
    outward unit normal and offset), face starts (f + 1), face vertices, edges (e, 2), local bounds
    (6) and, when known, volume and centroid. desc() gives the pxg_hull record over these arrays
    (keep this object alive while the record is used)."""

    def __init__(self, vertices, planes, face_starts, face_vertices, edges, bounds6, volume=None, centroid=None):
        self.vertices = np.ascontiguousarray(np.asarray(vertices, np.float32).reshape(-1, 3))
        self.planes = np.ascontiguousarray(np.asarray(planes, np.float32).reshape(-1, 4))
        self.face_starts = np.ascontiguousarray(np.asarray(face_starts, np.uint32))
        self.face_vertices = np.ascontiguousarray(np.asarray(face_vertices, np.uint32))
        self.edges = np.ascontiguousarray(np.asarray(edges, np.uint32).reshape(-1, 2))
        self.bounds6 = np.asarray(bounds6, np.float32)
        self.volume = volume
        self.centroid = centroid

    @property
    def n_faces(self):
        return len(self.planes)

    def face(self, f):
        return self.face_vertices[self.face_starts[f]:self.face_starts[f + 1]]

    def desc(self):
        d = HullDesc()
        d.vertices = self.vertices.ctypes.data
        d.n_vertices = len(self.vertices)
        d.n_faces = len(self.planes)
        d.face_planes = self.planes.ctypes.data
        d.face_starts = self.face_starts.ctypes.data
        d.face_vertices = self.face_vertices.ctypes.data if len(self.face_vertices) else None
        d.edges = self.edges.ctypes.data if len(self.edges) else None
        d.n_edges = len(self.edges)
        for i in range(3):
            d.local_min[i] = float(self.bounds6[i])
            d.local_max[i] = float(self.bounds6[3 + i])
        if self.centroid is None:
            self.volume, self.centroid = solid_volume_centroid(self.vertices, self.face_starts, self.face_vertices)
        for i in range(3):
            d.centroid[i] = float(self.centroid[i])
        d.volume = float(self.volume or 0.0)
        return d


def solid_volume_centroid(vertices, face_starts, face_vertices):
    """Volume and centroid of a closed convex polyhedron (faces counter-clockwise about their
    outward normals): signed tetrahedra from the vertex mean to each face's fan triangles, in
    float64. For hulls given without the builder's own centroid."""
    v = np.asarray(vertices, np.float64)
    o = v.mean(axis=0)
    vol = 0.0
    acc = np.zeros(3)
    for f in range(len(face_starts) - 1):
        idx = face_vertices[face_starts[f]:face_starts[f + 1]]
        for k in range(1, len(idx) - 1):
            a, b, c = v[idx[0]] - o, v[idx[k]] - o, v[idx[k + 1]] - o
            t = np.dot(a, np.cross(b, c)) / 6.0
            vol += t
            acc += t * (a + b + c) / 4.0
    cen = o + (acc / vol if vol > 0 else 0.0)
    return vol, cen.astype(np.float32)


def hull_table(hulls):
    """A ctypes array of pxg_hull records for a list of Hull objects."""
    arr = (HullDesc * max(1, len(hulls)))()
    for i, h in enumerate(hulls):
        arr[i] = h.desc()
    return arr



class MeshDesc(ctypes.Structure):
    """pxg_mesh_desc (include/pxg.h): a triangle soup's positions and index triples."""
    _fields_ = [("positions", ctypes.c_void_p), ("n_positions", ctypes.c_uint32), ("n_indices", ctypes.c_uint32),
                ("indices", ctypes.c_void_p)]


class MeshInfo(ctypes.Structure):
    """pxg_mesh_info (include/pxg.h): sizes and local bounds of a cooked mesh."""
    _fields_ = [("n_vertices", ctypes.c_uint32), ("n_triangles", ctypes.c_uint32), ("n_nodes", ctypes.c_uint32),
                ("non_manifold_edges", ctypes.c_uint32), ("local_min", ctypes.c_float * 3),
                ("local_max", ctypes.c_float * 3)]


class TriangleMesh:
    """A triangle soup for PXG_TRIANGLE_MESH shapes: positions (n, 3) and indices (3 per triangle),
    cooked by the library as TriangleMesh::from_triangles (geometry/triangle_mesh.h:75-170) does.
    desc() gives the pxg_mesh_desc record over these arrays (keep this object alive while the
    record is used)."""

    def __init__(self, positions, indices):
        self.positions = np.ascontiguousarray(np.asarray(positions, np.float32).reshape(-1, 3))
        self.indices = np.ascontiguousarray(np.asarray(indices, np.uint32).reshape(-1))

    def desc(self):
        d = MeshDesc()
        d.positions = self.positions.ctypes.data if len(self.positions) else None
        d.n_positions = len(self.positions)
        d.indices = self.indices.ctypes.data if len(self.indices) else None
        d.n_indices = len(self.indices)
        return d


def mesh_table(meshes):
    """A ctypes array of pxg_mesh_desc records for a list of TriangleMesh objects."""
    arr = (MeshDesc * max(1, len(meshes)))()
    for i, m in enumerate(meshes):
        arr[i] = m.desc()
    return arr


# ---- cloth and hair (include/pxg.h: pxg_cloth, pxg_strand_set) ----

PARTICLE_DTYPE = np.dtype([("position", "<f4", 3), ("inv_mass", "<f4"), ("prev_position", "<f4", 3),
                           ("reserved0", "<f4"), ("velocity", "<f4", 3), ("reserved1", "<f4")])
CONSTRAINT_DTYPE = np.dtype([("a", "<u4"), ("b", "<u4"), ("rest_length", "<f4"), ("lambda", "<f4")])
TETHER_DTYPE = np.dtype([("particle", "<u4"), ("anchor", "<u4"), ("max_length", "<f4"), ("reserved", "<u4")])
ANCHOR_DTYPE = np.dtype([("particle", "<u4"), ("has_body", "<u4"), ("body", "<u4"), ("generation", "<u4"),
                         ("local_position", "<f4", 3), ("world_position", "<f4", 3),
                         ("prev_world_position", "<f4", 3), ("reserved", "<f4")])
STRAND_DTYPE = np.dtype([("first", "<u4"), ("count", "<u4"), ("root_position", "<f4", 3),
                         ("root_orientation", "<f4", 4), ("prev_root_position", "<f4", 3),
                         ("prev_root_orientation", "<f4", 4)])
assert PARTICLE_DTYPE.itemsize == 48 and ANCHOR_DTYPE.itemsize == 56 and STRAND_DTYPE.itemsize == 64


class ClothParams(ctypes.Structure):
    """pxg_cloth_params (cloth::ClothParams, cloth/cloth.h:120-226) with the reference defaults."""
    _fields_ = [("stretch_compliance", ctypes.c_float), ("bend_compliance", ctypes.c_float),
                ("damping", ctypes.c_float), ("thickness", ctypes.c_float), ("friction", ctypes.c_float),
                ("gravity_scale", ctypes.c_float), ("max_velocity", ctypes.c_float),
                ("external_acceleration", ctypes.c_float * 3), ("wind", ctypes.c_float * 3),
                ("wind_turbulence", ctypes.c_float), ("air_drag", ctypes.c_float), ("air_lift", ctypes.c_float),
                ("self_collision", ctypes.c_uint32), ("self_distance", ctypes.c_float), ("tether_scale", ctypes.c_float),
                ("substeps", ctypes.c_uint32), ("iterations", ctypes.c_uint32), ("sleep_threshold", ctypes.c_float),
                ("sleep_time", ctypes.c_float), ("layer", ctypes.c_uint32)]

    def __init__(self, **kw):
        super().__init__()
        self.bend_compliance, self.damping, self.thickness, self.friction = 5e-5, 0.05, 0.03, 0.5
        self.gravity_scale, self.max_velocity, self.air_drag, self.air_lift = 1.0, 100.0, 0.02, 0.02
        self.tether_scale, self.sleep_threshold, self.sleep_time = 1.02, 0.05, 0.5
        for k, v in kw.items():
            if k in ("wind", "external_acceleration"):
                for i in range(3):
                    getattr(self, k)[i] = v[i]
            else:
                setattr(self, k, v)


class ClothDesc(ctypes.Structure):
    """pxg_cloth (cloth::Cloth, cloth/cloth.h:249-315) over caller-owned arrays."""
    _fields_ = [("particles", ctypes.c_void_p), ("n_particles", ctypes.c_uint32), ("columns", ctypes.c_uint32),
                ("rows", ctypes.c_uint32), ("stretch", ctypes.c_void_p), ("n_stretch", ctypes.c_uint32),
                ("bend", ctypes.c_void_p), ("n_bend", ctypes.c_uint32), ("stretch_batches", ctypes.c_void_p),
                ("n_stretch_batches", ctypes.c_uint32), ("bend_batches", ctypes.c_void_p),
                ("n_bend_batches", ctypes.c_uint32), ("tethers", ctypes.c_void_p), ("n_tethers", ctypes.c_uint32),
                ("anchors", ctypes.c_void_p), ("n_anchors", ctypes.c_uint32), ("triangles", ctypes.c_void_p),
                ("n_triangle_indices", ctypes.c_uint32), ("params", ClothParams), ("alive", ctypes.c_uint32),
                ("enabled", ctypes.c_uint32), ("awake", ctypes.c_uint32), ("sleep_timer", ctypes.c_float),
                ("time", ctypes.c_float), ("bounds_min", ctypes.c_float * 3), ("bounds_max", ctypes.c_float * 3)]


def _p(a):
    return a.ctypes.data if a is not None and len(a) else None


class Cloth:
    """A cloth's arrays (particles, stretch and bend constraints, batches as (begin, end) rows,
    tethers, anchors, triangles) with its grid size, parameters and sleep state. desc() gives the
    pxg_cloth record over them (keep the object alive while the record is used); read() fills them
    back from one."""

    def __init__(self, particles, stretch, bend, stretch_batches, bend_batches, tethers, anchors, triangles,
                 columns=0, rows=0, params=None, enabled=True, awake=True, sleep_timer=0.0, time=0.0, alive=True):
        self.particles = np.ascontiguousarray(particles, PARTICLE_DTYPE)
        self.stretch = np.ascontiguousarray(stretch, CONSTRAINT_DTYPE)
        self.bend = np.ascontiguousarray(bend, CONSTRAINT_DTYPE)
        self.stretch_batches = np.ascontiguousarray(np.asarray(stretch_batches, np.uint32).reshape(-1, 2))
        self.bend_batches = np.ascontiguousarray(np.asarray(bend_batches, np.uint32).reshape(-1, 2))
        self.tethers = np.ascontiguousarray(tethers, TETHER_DTYPE)
        self.anchors = np.ascontiguousarray(anchors, ANCHOR_DTYPE)
        self.triangles = np.ascontiguousarray(np.asarray(triangles, np.uint32).reshape(-1))
        self.columns, self.rows = columns, rows
        self.params = params if params is not None else ClothParams()
        self.enabled, self.awake, self.sleep_timer, self.time, self.alive = enabled, awake, sleep_timer, time, alive
        self.bounds = np.zeros(6, np.float32)

    def copy(self):
        c = Cloth(self.particles.copy(), self.stretch.copy(), self.bend.copy(), self.stretch_batches.copy(),
                  self.bend_batches.copy(), self.tethers.copy(), self.anchors.copy(), self.triangles.copy(),
                  self.columns, self.rows, ClothParams.from_buffer_copy(self.params), self.enabled, self.awake,
                  self.sleep_timer, self.time, self.alive)
        c.bounds = self.bounds.copy()
        return c

    def desc(self):
        d = ClothDesc()
        d.particles, d.n_particles = _p(self.particles), len(self.particles)
        d.columns, d.rows = self.columns, self.rows
        d.stretch, d.n_stretch = _p(self.stretch), len(self.stretch)
        d.bend, d.n_bend = _p(self.bend), len(self.bend)
        d.stretch_batches, d.n_stretch_batches = _p(self.stretch_batches), len(self.stretch_batches)
        d.bend_batches, d.n_bend_batches = _p(self.bend_batches), len(self.bend_batches)
        d.tethers, d.n_tethers = _p(self.tethers), len(self.tethers)
        d.anchors, d.n_anchors = _p(self.anchors), len(self.anchors)
        d.triangles, d.n_triangle_indices = _p(self.triangles), len(self.triangles)
        d.params = self.params
        d.alive, d.enabled, d.awake = int(self.alive), int(self.enabled), int(self.awake)
        d.sleep_timer, d.time = self.sleep_timer, self.time
        for i in range(3):
            d.bounds_min[i], d.bounds_max[i] = float(self.bounds[i]), float(self.bounds[3 + i])
        return d

    def take_state(self, d):
        """The sleep state, time and bounds of a pxg_cloth record filled by a read-back."""
        self.awake, self.sleep_timer, self.time = bool(d.awake), float(d.sleep_timer), float(d.time)
        self.bounds = np.array(list(d.bounds_min) + list(d.bounds_max), np.float32)


def cloth_table(cloths):
    arr = (ClothDesc * max(1, len(cloths)))()
    for i, c in enumerate(cloths):
        arr[i] = c.desc()
    return arr


class StrandParams(ctypes.Structure):
    """pxg_strand_params (hair::StrandParams, hair/strand.h:81-180) with the reference defaults."""
    _fields_ = [("stretch_compliance", ctypes.c_float), ("global_shape", ctypes.c_float),
                ("global_shape_falloff", ctypes.c_float), ("local_shape", ctypes.c_float), ("damping", ctypes.c_float),
                ("thickness", ctypes.c_float), ("friction", ctypes.c_float), ("gravity_scale", ctypes.c_float),
                ("max_velocity", ctypes.c_float), ("external_acceleration", ctypes.c_float * 3),
                ("wind", ctypes.c_float * 3), ("wind_turbulence", ctypes.c_float), ("air_drag", ctypes.c_float),
                ("substeps", ctypes.c_uint32), ("iterations", ctypes.c_uint32), ("sleep_threshold", ctypes.c_float),
                ("sleep_time", ctypes.c_float), ("layer", ctypes.c_uint32)]

    def __init__(self, **kw):
        super().__init__()
        self.global_shape, self.global_shape_falloff, self.local_shape, self.damping = 0.22, 1.6, 0.35, 0.08
        self.thickness, self.friction, self.gravity_scale, self.max_velocity = 0.01, 0.4, 1.0, 50.0
        self.wind_turbulence, self.air_drag, self.sleep_threshold, self.sleep_time = 0.1, 0.6, 0.02, 0.6
        for k, v in kw.items():
            if k in ("wind", "external_acceleration"):
                for i in range(3):
                    getattr(self, k)[i] = v[i]
            else:
                setattr(self, k, v)


class StrandSetDesc(ctypes.Structure):
    """pxg_strand_set (hair::StrandSet, hair/strand.h:187-240) over caller-owned arrays."""
    _fields_ = [("particles", ctypes.c_void_p), ("n_particles", ctypes.c_uint32), ("stretch", ctypes.c_void_p),
                ("n_stretch", ctypes.c_uint32), ("stretch_batches", ctypes.c_void_p),
                ("n_stretch_batches", ctypes.c_uint32), ("rest_local", ctypes.c_void_p),
                ("rest_offset", ctypes.c_void_p), ("strands", ctypes.c_void_p), ("n_strands", ctypes.c_uint32),
                ("collide_bodies", ctypes.c_void_p), ("n_collide_bodies", ctypes.c_uint32),
                ("params", StrandParams), ("alive", ctypes.c_uint32), ("enabled", ctypes.c_uint32),
                ("awake", ctypes.c_uint32), ("sleep_timer", ctypes.c_float), ("time", ctypes.c_float),
                ("bounds_min", ctypes.c_float * 3), ("bounds_max", ctypes.c_float * 3)]


class StrandSet:
    """A groom's arrays (particles, stretch constraints and batches, rest poses, strands, collide
    bodies as (index, generation) rows) with its parameters and sleep state."""

    def __init__(self, particles, stretch, stretch_batches, rest_local, rest_offset, strands, collide_bodies=(),
                 params=None, enabled=True, awake=True, sleep_timer=0.0, time=0.0, alive=True):
        self.particles = np.ascontiguousarray(particles, PARTICLE_DTYPE)
        self.stretch = np.ascontiguousarray(stretch, CONSTRAINT_DTYPE)
        self.stretch_batches = np.ascontiguousarray(np.asarray(stretch_batches, np.uint32).reshape(-1, 2))
        self.rest_local = np.ascontiguousarray(np.asarray(rest_local, np.float32).reshape(-1, 3))
        self.rest_offset = np.ascontiguousarray(np.asarray(rest_offset, np.float32).reshape(-1, 3))
        self.strands = np.ascontiguousarray(strands, STRAND_DTYPE)
        self.collide_bodies = np.ascontiguousarray(np.asarray(collide_bodies, np.uint32).reshape(-1, 2))
        self.params = params if params is not None else StrandParams()
        self.enabled, self.awake, self.sleep_timer, self.time, self.alive = enabled, awake, sleep_timer, time, alive
        self.bounds = np.zeros(6, np.float32)

    def copy(self):
        s = StrandSet(self.particles.copy(), self.stretch.copy(), self.stretch_batches.copy(), self.rest_local.copy(),
                      self.rest_offset.copy(), self.strands.copy(), self.collide_bodies.copy(),
                      StrandParams.from_buffer_copy(self.params), self.enabled, self.awake, self.sleep_timer,
                      self.time, self.alive)
        s.bounds = self.bounds.copy()
        return s

    def desc(self):
        d = StrandSetDesc()
        d.particles, d.n_particles = _p(self.particles), len(self.particles)
        d.stretch, d.n_stretch = _p(self.stretch), len(self.stretch)
        d.stretch_batches, d.n_stretch_batches = _p(self.stretch_batches), len(self.stretch_batches)
        d.rest_local, d.rest_offset = _p(self.rest_local), _p(self.rest_offset)
        d.strands, d.n_strands = _p(self.strands), len(self.strands)
        d.collide_bodies, d.n_collide_bodies = _p(self.collide_bodies), len(self.collide_bodies)
        d.params = self.params
        d.alive, d.enabled, d.awake = int(self.alive), int(self.enabled), int(self.awake)
        d.sleep_timer, d.time = self.sleep_timer, self.time
        for i in range(3):
            d.bounds_min[i], d.bounds_max[i] = float(self.bounds[i]), float(self.bounds[3 + i])
        return d

    def take_state(self, d):
        self.awake, self.sleep_timer, self.time = bool(d.awake), float(d.sleep_timer), float(d.time)
        self.bounds = np.array(list(d.bounds_min) + list(d.bounds_max), np.float32)


def strand_table(sets):
    arr = (StrandSetDesc * max(1, len(sets)))()
    for i, s in enumerate(sets):
        arr[i] = s.desc()
    return arr
