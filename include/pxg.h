/*
 * pxg.h -- C ABI of the MI355X rigid-body contact step.
 *
 * This is the drop-in boundary for physxcoopa's `PhysicsWorld::step_fixed()` hot path:
 * broadphase pair finding, narrowphase (box/sphere/capsule), contact-graph colouring, the
 * colour-by-colour sequential-impulse solver, integration, islands and sleep. Host code keeps
 * the reference's body/shape/material model; everything here is plain C: opaque handles,
 * plain structs, pointers and sizes. Every function returns an int status (PXG_OK == 0) and
 * never throws; pxg_last_error() returns the message of the last failure on this thread.
 *
 * Conventions (same as the reference):
 *   - Z-up, metres, seconds, kg. Gravity default (0, 0, -9.81).
 *   - Quaternions are stored x, y, z, w (glm's storage order).
 *   - Matrices are 3x3 column-major (glm::mat3 layout): m[c*3 + r].
 *   - Body type order Static, Kinematic, Dynamic  (physxcoopa/dynamics/body.h:19-23).
 *   - Shape type order Sphere, Box, Capsule        (physxcoopa/collision/shape.h:37-43).
 *   - Combine mode order Average, Minimum, Maximum, Multiply
 *                                                  (physxcoopa/dynamics/physics_material.h:28-33).
 *   - Body index == position in the body table; shape slot == position in the shape table.
 *
 * One host thread per handle. Each world handle owns one HIP stream on the current device.
 */
#ifndef PXG_H
#define PXG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PXG_OK 0
#define PXG_ERR_INVALID 1   /* bad argument or handle */
#define PXG_ERR_CAPACITY 2  /* a fixed-capacity buffer would overflow */
#define PXG_ERR_HIP 3       /* a HIP runtime call failed */
#define PXG_ERR_STATE 4     /* call not valid in the current state */
#define PXG_ERR_DEVICE 5    /* a kernel reported an internal error (bounded wait expired, ...) */

enum pxg_body_type { PXG_STATIC = 0, PXG_KINEMATIC = 1, PXG_DYNAMIC = 2 };
/* How a renderer blends a body's previous and current pose between fixed substeps
 * (dynamics::Interpolation, dynamics/body.h:26-33; Body::interpolation defaults to Interpolate,
 * body.h:118). Carried with the body; the step does not read it. */
enum pxg_interpolation { PXG_INTERPOLATE = 0, PXG_INTERPOLATE_NONE = 1 };
enum pxg_shape_type { PXG_SPHERE = 0, PXG_BOX = 1, PXG_CAPSULE = 2, PXG_CONVEX_HULL = 3, PXG_TRIANGLE_MESH = 4 };
enum pxg_combine { PXG_AVERAGE = 0, PXG_MINIMUM = 1, PXG_MAXIMUM = 2, PXG_MULTIPLY = 3 };
enum pxg_phase { PXG_PHASE_FORCES_COLLIDE = 0, PXG_PHASE_SOLVE_INTEGRATE = 1 };

/* Body constraint bits (physxcoopa/dynamics/body.h:37-47). */
#define PXG_FREEZE_POS_X (1u << 0)
#define PXG_FREEZE_POS_Y (1u << 1)
#define PXG_FREEZE_POS_Z (1u << 2)
#define PXG_FREEZE_ROT_X (1u << 3)
#define PXG_FREEZE_ROT_Y (1u << 4)
#define PXG_FREEZE_ROT_Z (1u << 5)

/*
 * Solver / broadphase tunables. Replaces util::PhysicsConfig (physxcoopa/util/config.h:22-145)
 * plus the world's gravity (world.h:2099) and the GPU capacities.
 */
typedef struct pxg_config {
    uint32_t velocity_iterations;     /* 16 */
    uint32_t relax_iterations;        /* 2  */
    uint32_t position_iterations;     /* 4  */
    float position_correction;        /* 0.2 */
    float linear_slop;                /* 0.005 */
    float max_linear_correction;      /* 0.2 */
    float restitution_threshold;      /* 0.5 m/s */
    float sleep_linear;               /* 0.01 m/s */
    float sleep_angular;              /* 0.02 rad/s */
    float sleep_time;                 /* 0.5 s */
    float aabb_margin;                /* 0.1 m */
    uint32_t max_substeps;            /* 8 */
    float fixed_dt;                   /* 1/60 s */
    float max_linear_velocity;        /* 200 m/s */
    uint32_t use_exact_quaternion_integration; /* 0 */
    float gravity[3];                 /* (0, 0, -9.81) */
    /* Residual mode: when residual_mode != 0 the velocity pass repeats until the velocity
     * residual (max over active contact rows, m/s) is <= residual_tolerance or
     * residual_max_iterations passes have run; relax and position passes are unchanged. */
    uint32_t residual_mode;           /* 0 */
    float residual_tolerance;         /* 1e-4 */
    uint32_t residual_max_iterations; /* 256 */
    /* Broadphase: shapes whose fat AABB is wider than this in x or y are paired by a
     * brute-force path instead of the x-y grid (a ground slab, say). */
    float large_extent;               /* 16 m */
    /* Capacities of the device buffers. 0 = derive from the first upload. */
    uint32_t max_bodies;
    uint32_t max_shapes;
    uint32_t max_pairs;
    uint32_t max_manifolds;
    /* Learned warm start: 1 = predict initial impulses for contacts with no cached impulse. */
    uint32_t learned_warm_start;      /* 0 */
    /* Solve order. PXG_ORDER_REFERENCE: manifolds that share an awake dynamic body are solved in
     * the reference's manifold order (pair order, broadphase/pair_cache.h:23-52), so the result
     * is the reference's sequential Gauss-Seidel result; manifolds that share no such body run in
     * parallel (colours = longest chains of that order). PXG_ORDER_COLOURED: greedy colouring,
     * fewer colours (faster), a different but equally valid sequential order. PXG_ORDER_TILED:
     * greedy colouring within spatial tiles of the awake bodies (manifolds crossing tiles in a
     * later phase, grouped by connected components), each tile solved by one workgroup in LDS:
     * the fastest order for large piles;
     * worlds with joints use PXG_ORDER_COLOURED instead. */
    uint32_t solve_order;             /* PXG_ORDER_REFERENCE */
    /* Residual mode: the residual is evaluated after every residual_check_interval-th velocity
     * pass (and after the last one the cap allows); the passes between run without the residual
     * sweep. 0 and 1 both mean every pass. */
    uint32_t residual_check_interval; /* 1 */
    /* Cloth and hair (util/config.h:139-144): XPBD substeps per 1/60 s and constraint iterations
     * per substep, unless a cloth or groom overrides them. */
    uint32_t cloth_substeps;          /* 4 */
    uint32_t cloth_iterations;        /* 1 */
    uint32_t reserved[3];
} pxg_config;

enum pxg_solve_order { PXG_ORDER_REFERENCE = 0, PXG_ORDER_COLOURED = 1, PXG_ORDER_TILED = 2 };

/* Full dynamic state of one body (physxcoopa/dynamics/body.h:85-216). */
typedef struct pxg_body {
    uint32_t type;          /* pxg_body_type */
    uint32_t alive;         /* 0 = free slot (ignored everywhere) */
    uint32_t awake;
    uint32_t use_gravity;
    uint32_t constraints;   /* PXG_FREEZE_* bits */
    uint32_t generation;    /* BodyId generation: cloth anchors and groom collide bodies name a body
                             * by (index, generation) and follow it only while both match */
    uint32_t interpolation; /* pxg_interpolation; 0 = PXG_INTERPOLATE, the reference's default */
    uint32_t reserved1;
    float position[3];
    float sleep_timer;
    float orientation[4];   /* x, y, z, w */
    float linear_velocity[3];
    float mass;
    float angular_velocity[3];
    float inv_mass;
    float force[3];
    float linear_drag;
    float torque[3];
    float angular_drag;
    float inv_inertia_local[3];
    float reserved2;
    float inv_inertia_world[9]; /* column-major; recomputed by the library */
    float prev_position[3];
    float reserved3;
    float prev_orientation[4];
} pxg_body;

/* One collision shape slot (physxcoopa/collision/shape.h:54-157) with its material
 * (physxcoopa/dynamics/physics_material.h:40-58) and layer. */
typedef struct pxg_shape {
    uint32_t owner;          /* body index */
    uint32_t type;           /* pxg_shape_type */
    uint32_t enabled;        /* 0 = no broadphase proxy (body without collider) */
    uint32_t is_trigger;
    uint32_t layer;          /* 0..31 */
    int32_t capsule_axis;    /* 0 = X, 1 = Y, 2 = Z; PXG_CONVEX_HULL: the hull's index in pxg_set_hulls' table;
                                PXG_TRIANGLE_MESH: the mesh's index in pxg_set_meshes' table */
    uint32_t friction_combine;
    uint32_t restitution_combine;
    float local_center[3];
    float radius;            /* sphere; PXG_CONVEX_HULL: the uniform scale (Shape::hull_scale, shape.h:98);
                                PXG_TRIANGLE_MESH: the uniform scale (Shape::mesh_scale, shape.h:91) */
    float local_rotation[4]; /* x, y, z, w */
    float half_extents[3];   /* box */
    float capsule_radius;
    float capsule_half_height;
    float dynamic_friction;  /* 0.6 default material */
    float static_friction;   /* carried, unused by the solver (as in the reference) */
    float restitution;       /* 0.0 default material */
} pxg_shape;

/* One contact point (physxcoopa/collision/manifold.h:27-52). */
typedef struct pxg_contact {
    float position[3];
    float penetration;
    float separation;
    float normal_impulse;
    float tangent_impulse[2];
    uint32_t feature_id;
    float restitution_bias;
    float reserved[2];
} pxg_contact;

/* One contact manifold (physxcoopa/collision/manifold.h:64-97). */
typedef struct pxg_manifold {
    uint32_t a, b;            /* body indices; normal points a -> b */
    uint32_t shape_a, shape_b;/* shape slots (the broadphase pair) */
    float normal[3];
    float friction;
    float restitution;
    uint32_t count;
    uint32_t is_trigger;
    uint32_t valid;
    pxg_contact points[4];
} pxg_manifold;

/* Per-step counters. */
typedef struct pxg_stats {
    uint32_t pairs;              /* broadphase pairs */
    uint32_t manifolds;          /* valid manifolds, triggers included */
    uint32_t points;             /* contact points in valid manifolds */
    uint32_t colours;            /* colours used by the solver this step */
    uint32_t active_manifolds;   /* manifolds the solver ran (non-trigger, an awake dynamic side) */
    uint32_t velocity_passes;    /* velocity passes run (== velocity_iterations unless residual mode) */
    uint32_t awake_bodies;
    uint32_t learned_predictions;/* warm-start misses filled by the learned predictor */
    float residual;              /* velocity residual after the last velocity pass (m/s) */
    uint32_t tiled_workgroups;   /* PXG_ORDER_TILED, the last solve: its workgroups (bits 0-7), how
                                    many held their tile in LDS (bits 8-15), and the most LDS a
                                    tile needed, in KiB (bits 16-31) */
    uint64_t steps;              /* fixed steps run since creation */
} pxg_stats;

/* Joint types (physxcoopa/dynamics/joint.h:57-61). */
enum pxg_joint_type { PXG_HINGE = 0, PXG_BALL = 1, PXG_CONE_TWIST = 2 };

/* One joint (physxcoopa/dynamics/joint.h:74-135). Anchors and axes are in each body's own frame.
 * A hinge locks two rotational DOFs (optional angle limit [min_angle, max_angle]); a ball joint
 * holds only the anchors together; a cone-twist adds a swing cone (swing_limit >= 0) and a twist
 * range (twist_min <= twist_max). Limits are measured from rest_relative_rotation =
 * inverse(qa) * qb at creation (world.h:415). The impulse accumulators are solver state: the
 * point and hinge-axis ones carry over between steps (the joint's warm start). */
typedef struct pxg_joint {
    uint32_t type;               /* pxg_joint_type */
    uint32_t body_a, body_b;     /* body indices */
    uint32_t valid;              /* 0: ignored by every pass (removed joint, or a removed body) */
    uint32_t enabled;
    uint32_t use_limits;         /* hinge angle limit on */
    uint32_t collide_connected;  /* 0: body_a and body_b generate no contacts with each other */
    uint32_t reserved0;
    float local_anchor_a[3];
    float min_angle;             /* radians */
    float local_anchor_b[3];
    float max_angle;
    float local_axis_a[3];       /* hinge axis / twist axis (normalised where read) */
    float swing_limit;           /* cone half-angle; negative: swing free */
    float local_axis_b[3];
    float twist_min;             /* twist_min > twist_max: twist free */
    float rest_relative_rotation[4];  /* x, y, z, w */
    float twist_max;
    float limit_impulse;         /* hinge limit / cone twist limit accumulator */
    float swing_impulse;
    float reserved1;
    float point_impulse[3];
    float reserved2;
    float axis_impulse[2];
    float reserved3[2];
} pxg_joint;

/* One contact event (physxcoopa/collision/contact_event.h:29-36): a shape pair's enter / stay /
 * exit transition of one step, for a solid contact or a trigger overlap. a and b are the owner
 * body indices of the pair's shapes (lower shape slot first). */
enum pxg_event_type { PXG_EVENT_ENTER = 0, PXG_EVENT_STAY = 1, PXG_EVENT_EXIT = 2 };
typedef struct pxg_contact_event {
    uint32_t a, b;
    uint32_t type;        /* pxg_event_type */
    uint32_t is_trigger;
} pxg_contact_event;

/* Scene queries (query/queries.h:56-98, geometry/ray.h:32-35). A ray starts at origin and runs
 * along direction (unit length) up to max_distance (FLT_MAX: unbounded). */
typedef struct pxg_ray {
    float origin[3];
    float max_distance;
    float direction[3];
    float reserved;
} pxg_ray;
#define PXG_NO_BODY 0xFFFFFFFFu
/* QueryFilter without its host-side predicate: bit shape.layer of layer_mask must be set, trigger
 * shapes are skipped unless include_triggers, the shapes of body ignore_body are skipped. A null
 * filter passes every shape. */
typedef struct pxg_query_filter {
    uint32_t layer_mask;       /* ~0u: every layer */
    uint32_t include_triggers; /* 1 */
    uint32_t ignore_body;      /* PXG_NO_BODY: none */
    uint32_t reserved;
} pxg_query_filter;
/* RaycastHit (query/queries.h:56-76): hit = 0 when nothing was hit (the rest is then zero);
 * started_inside: shape casts only, the cast shape already overlapped the target at its start. */
typedef struct pxg_raycast_hit {
    float point[3];
    float distance;
    float normal[3];
    uint32_t hit;
    uint32_t body;
    uint32_t shape_index;
    uint32_t started_inside;
    uint32_t reserved;
} pxg_raycast_hit;
/* One shape cast (shape_cast(), world.h:877-909): the shape (sphere, box or capsule; its local centre,
 * local rotation, capsule axis and dimensions are used, owner / material / layer are not) posed at
 * position / orientation, swept along direction (normalised by the library) up to max_distance. */
typedef struct pxg_cast_query {
    pxg_shape shape;
    float position[3];
    float max_distance;
    float orientation[4];
    float direction[3];
    float reserved;
} pxg_cast_query;

/* The query shape of overlap_sphere / overlap_box / overlap_capsule (world.h:1002-1060). */
typedef struct pxg_overlap_query {
    uint32_t type;          /* PXG_SPHERE, PXG_BOX or PXG_CAPSULE */
    float radius;           /* sphere, capsule */
    float center[3];        /* sphere, box */
    float half_extents[3];  /* box */
    float orientation[4];   /* box: x, y, z, w */
    float a[3], b[3];       /* capsule end points */
} pxg_overlap_query;

/* One overlap found by compute_penetration() (query::Penetration, query/queries.h:111-117): moving
 * the query shape by normal * depth separates it from the target shape; normal is unit and points
 * from the target toward the query shape; point is a representative contact point on the target. */
typedef struct pxg_penetration {
    float normal[3];
    float depth;
    float point[3];
    uint32_t body;
    uint32_t shape_index;
    uint32_t reserved;
} pxg_penetration;

typedef struct pxg_world pxg_world;

/* --- configuration and lifetime --- */
void pxg_config_default(pxg_config* cfg);
int pxg_create(const pxg_config* cfg, pxg_world** out);          /* PhysicsWorld(config), world.h:81 */
int pxg_destroy(pxg_world* w);
int pxg_get_config(pxg_world* w, pxg_config* out);                /* config(), world.h:1207 */
int pxg_set_config(pxg_world* w, const pxg_config* cfg);          /* capacities must not change */
int pxg_set_gravity(pxg_world* w, const float g[3]);              /* set_gravity(), world.h:1144 */

/* A built convex hull (geometry::ConvexHull, geometry/convex_hull.h:38-94), as the reference's
 * build_convex_hull() leaves it: vertices, merged polygon faces (outward unit normal and plane
 * offset d, vertex lists counter-clockwise about the normal), unique edges (vertex index pairs,
 * lower first) and the local bounds. Caller-owned arrays, copied by pxg_set_hulls. */
typedef struct pxg_hull {
    const float* vertices;         /* 3 per vertex */
    uint32_t n_vertices;           /* at most 64 (the builder's default cap) */
    uint32_t n_faces;
    const float* face_planes;      /* 4 per face: normal xyz, d */
    const uint32_t* face_starts;   /* n_faces + 1: face f spans [face_starts[f], face_starts[f + 1]) */
    const uint32_t* face_vertices; /* at most 8 per face on the GPU */
    const uint32_t* edges;         /* 2 per edge */
    uint32_t n_edges;
    float local_min[3];
    float local_max[3];
    float centroid[3];             /* the solid's centroid (what a hull resting on a mesh is measured from) */
    float volume;
} pxg_hull;

/* A triangle soup, as TriangleMesh::from_triangles (geometry/triangle_mesh.h:75-170) takes it:
 * positions and index triples. Cooking welds positions closer than 1e-4 of the bounds' diagonal,
 * drops the triangles the weld collapses or leaves without area, and computes unit face normals,
 * edge adjacency and the triangle tree (geometry/mesh_bvh.h). Caller-owned arrays. */
typedef struct pxg_mesh_desc {
    const float* positions;        /* 3 per position */
    uint32_t n_positions;
    uint32_t n_indices;            /* 3 per triangle; a trailing partial triangle is ignored */
    const uint32_t* indices;       /* each < n_positions */
} pxg_mesh_desc;

/* Sizes and bounds of a cooked mesh. */
typedef struct pxg_mesh_info {
    uint32_t n_vertices;           /* after welding (at most n_positions) */
    uint32_t n_triangles;          /* after dropping degenerate triangles (at most n_indices / 3) */
    uint32_t n_nodes;              /* triangle tree nodes */
    uint32_t non_manifold_edges;   /* edges shared by more than two triangles (left unlinked) */
    float local_min[3];            /* bounds of every welded vertex */
    float local_max[3];
} pxg_mesh_info;

/* --- cloth and hair (XPBD, stepped once per frame by pxg_step: world.h:1241-1256) --- */
/* cloth::ClothParticle (cloth/cloth.h:52-57), also a strand particle. */
typedef struct pxg_particle {
    float position[3];
    float inv_mass;                /* 0 = pinned */
    float prev_position[3];
    float reserved0;
    float velocity[3];
    float reserved1;
} pxg_particle;

/* cloth::ClothConstraint (cloth/cloth.h:61-66): an XPBD distance constraint. */
typedef struct pxg_distance_constraint {
    uint32_t a, b;
    float rest_length;
    float lambda;                  /* the last substep's multiplier (reset every substep) */
} pxg_distance_constraint;

/* cloth::ClothTether (cloth/cloth.h:75-79). */
typedef struct pxg_cloth_tether {
    uint32_t particle, anchor;
    float max_length;
    uint32_t reserved;
} pxg_cloth_tether;

/* cloth::ClothAnchor (cloth/cloth.h:89-106): a pinned particle following a body, or a static pin. */
typedef struct pxg_cloth_anchor {
    uint32_t particle;
    uint32_t has_body;             /* 0: a static pin at world_position */
    uint32_t body, generation;     /* the body followed (dynamics::BodyId); stale = static pin */
    float local_position[3];
    float world_position[3];
    float prev_world_position[3];
    float reserved;
} pxg_cloth_anchor;

/* cloth::ClothParams (cloth/cloth.h:120-226). */
typedef struct pxg_cloth_params {
    float stretch_compliance;      /* 0 */
    float bend_compliance;         /* 5e-5 */
    float damping;                 /* 0.05 */
    float thickness;               /* 0.03 */
    float friction;                /* 0.5 */
    float gravity_scale;           /* 1 */
    float max_velocity;            /* 100 */
    float external_acceleration[3];
    float wind[3];
    float wind_turbulence;         /* 0 */
    float air_drag;                /* 0.02 */
    float air_lift;                /* 0.02 */
    uint32_t self_collision;       /* 0 */
    float self_distance;           /* 0: the builder's 0.6 x the particle spacing */
    float tether_scale;            /* 1.02 */
    uint32_t substeps;             /* 0: the world's */
    uint32_t iterations;           /* 0: the world's */
    float sleep_threshold;         /* 0.05 */
    float sleep_time;              /* 0.5 */
    uint32_t layer;
} pxg_cloth_params;

/* cloth::Cloth (cloth/cloth.h:249-315) over caller-owned arrays: particles, stretch and bend
 * constraints with their particle-disjoint batches ([begin, end) pairs), tethers, anchors, the
 * triangles aerodynamics uses, the grid size (columns 0: not a grid; self-collision skips grid
 * neighbours by it) and the sleep state. alive 0 is an empty slot. */
typedef struct pxg_cloth {
    pxg_particle* particles;
    uint32_t n_particles;
    uint32_t columns, rows;
    pxg_distance_constraint* stretch;
    uint32_t n_stretch;
    pxg_distance_constraint* bend;
    uint32_t n_bend;
    const uint32_t* stretch_batches;
    uint32_t n_stretch_batches;
    const uint32_t* bend_batches;
    uint32_t n_bend_batches;
    const pxg_cloth_tether* tethers;
    uint32_t n_tethers;
    pxg_cloth_anchor* anchors;
    uint32_t n_anchors;
    const uint32_t* triangles;
    uint32_t n_triangle_indices;
    pxg_cloth_params params;
    uint32_t alive, enabled, awake;
    float sleep_timer, time;
    float bounds_min[3], bounds_max[3];
} pxg_cloth;

/* hair::Strand (hair/strand.h:40-72): a chain of particles from a driven root. */
typedef struct pxg_strand {
    uint32_t first, count;
    float root_position[3];
    float root_orientation[4];     /* x, y, z, w */
    float prev_root_position[3];
    float prev_root_orientation[4];
} pxg_strand;

/* hair::StrandParams (hair/strand.h:81-180). */
typedef struct pxg_strand_params {
    float stretch_compliance;      /* 0 */
    float global_shape;            /* 0.22 */
    float global_shape_falloff;    /* 1.6 */
    float local_shape;             /* 0.35 */
    float damping;                 /* 0.08 */
    float thickness;               /* 0.01 */
    float friction;                /* 0.4 */
    float gravity_scale;           /* 1 */
    float max_velocity;            /* 50 */
    float external_acceleration[3];
    float wind[3];
    float wind_turbulence;         /* 0.1 */
    float air_drag;                /* 0.6 */
    uint32_t substeps;             /* 0: the world's */
    uint32_t iterations;           /* 0: the world's */
    float sleep_threshold;         /* 0.02 */
    float sleep_time;              /* 0.6 */
    uint32_t layer;
} pxg_strand_params;

/* hair::StrandSet (hair/strand.h:187-240) over caller-owned arrays. collide_bodies: (index,
 * generation) pairs; when given, the only bodies the groom collides with. alive 0 is an empty slot. */
typedef struct pxg_strand_set {
    pxg_particle* particles;
    uint32_t n_particles;
    pxg_distance_constraint* stretch;
    uint32_t n_stretch;
    const uint32_t* stretch_batches;
    uint32_t n_stretch_batches;
    const float* rest_local;       /* 3 per particle: rest position in its strand's root frame */
    const float* rest_offset;      /* 3 per particle: rest offset from the previous particle */
    pxg_strand* strands;
    uint32_t n_strands;
    const uint32_t* collide_bodies;
    uint32_t n_collide_bodies;
    pxg_strand_params params;
    uint32_t alive, enabled, awake;
    float sleep_timer, time;
    float bounds_min[3], bounds_max[3];
} pxg_strand_set;

/* --- state upload / download (caller-owned host buffers are copied) --- */
/* Replaces the body table (add_body/remove_body, world.h:95-143). World inverse inertia is
 * recomputed from orientation exactly as add_body() does. */
int pxg_set_bodies(pxg_world* w, const pxg_body* bodies, uint32_t n);
int pxg_get_bodies(pxg_world* w, pxg_body* out, uint32_t n);
/* Writes records[k] into body indices[k] for k < n and leaves every other body as it is: what
 * edits through get_body() (world.h:347) need to reach the GPU, without the whole table. World
 * inverse inertia of the written bodies is recomputed as pxg_set_bodies does. Indices must be
 * distinct and below the body count, and a body's type may not change (PXG_ERR_STATE: use
 * pxg_set_bodies); allowed between the two step phases like pxg_set_bodies. */
int pxg_set_bodies_subset(pxg_world* w, const uint32_t* indices, const pxg_body* records, uint32_t n);

/* What a renderer's transform write-back reads of a body (system/physics_system.h:1058-1075:
 * type, interpolation mode, previous and current pose), 64 bytes per body. */
typedef struct pxg_pose {
    float position[3];
    uint32_t type;            /* pxg_body_type */
    float orientation[4];     /* x, y, z, w */
    float prev_position[3];
    uint32_t interpolation;   /* pxg_interpolation */
    float prev_orientation[4];
} pxg_pose;
/* The poses of bodies [0, n): one download of 64 bytes per body instead of the full records. */
int pxg_get_poses(pxg_world* w, pxg_pose* out, uint32_t n);
/* Replaces the shape table and creates each enabled shape's broadphase proxy with
 * fat = tight.expand(margin) at its owner's current pose (create_shape_proxy_, world.h:1369-1382). */
int pxg_set_shapes(pxg_world* w, const pxg_shape* shapes, uint32_t n);
/* Replaces the table of convex hulls that PXG_CONVEX_HULL shapes index (the hulls a
 * ConvexCollider's Shape::hull points at, shape.h:93-98); upload before the shapes that use them.
 * Hull narrowphase: collision/sat_hull.h (hull against hull, box, sphere and capsule; SAT, no
 * speculative contacts). A hull with more than 64 vertices or a face of more than 8 vertices is
 * refused (PXG_ERR_INVALID). */
int pxg_set_hulls(pxg_world* w, const pxg_hull* hulls, uint32_t n);
/* Cooks one mesh on the host (no device use) as pxg_set_meshes does: info receives the sizes;
 * each non-null array receives the cooked data -- vertices (3 floats per vertex; 3 * n_positions
 * floats always suffice), indices (3 per triangle; n_indices suffice), normals (3 floats per
 * triangle; n_indices suffice) and adjacency (per triangle the neighbour across edges v0-v1,
 * v1-v2, v2-v0, 0xFFFFFFFF for none; n_indices suffice). PXG_ERR_INVALID for a null array with
 * n > 0 or an index out of range. TriangleMesh::from_triangles, triangle_mesh.h:75-170. */
int pxg_cook_mesh(const pxg_mesh_desc* mesh, pxg_mesh_info* info, float* vertices, uint32_t* indices, float* normals,
                  uint32_t* adjacency);
/* Replaces the table of triangle meshes that PXG_TRIANGLE_MESH shapes index (the meshes a
 * MeshCollider's Shape::mesh points at, shape.h:84-91), cooking each as pxg_cook_mesh does;
 * upload before the shapes that use them. A mesh shape belongs to a static or kinematic body.
 * Mesh narrowphase: collision/mesh_contact.h (sphere, capsule, box and convex hull against the
 * mesh's triangles, internal-edge correction, embedded recovery; no mesh against mesh, no
 * speculative contacts). A mesh that cooks to no triangle is refused (PXG_ERR_INVALID). */
int pxg_set_meshes(pxg_world* w, const pxg_mesh_desc* meshes, uint32_t n);
/* Replaces the cloth table (add_cloth / remove_cloth / get_cloth, world.h:552-620): slot i is
 * cloths[i]; a slot with alive 0 or no particle is empty. Cloth is advanced by pxg_step once per
 * frame after the rigid substeps (step_cloths_, world.h:1749-1847), collides with the world's
 * shapes one-way and is part of pxg_state_hash. Batches must be particle-disjoint ranges of their
 * constraint arrays and every index in range (PXG_ERR_INVALID otherwise); at most 8192 particles
 * per cloth. */
int pxg_set_cloths(pxg_world* w, const pxg_cloth* cloths, uint32_t n);
/* Reads cloth slot `index` back into out's arrays (which must have the uploaded sizes; any may
 * be null): particles, constraints' multipliers, anchors' world positions, and the sleep state,
 * time and bounds. */
int pxg_get_cloth(pxg_world* w, uint32_t index, pxg_cloth* out);
/* Replaces the hair table (add_strands / remove_strands / get_strands, world.h:623-700), stepped
 * by pxg_step after the cloths (step_strands_, world.h:1861-1957). Not hashed (world.h:1334-1359). */
int pxg_set_strand_sets(pxg_world* w, const pxg_strand_set* sets, uint32_t n);
int pxg_get_strand_set(pxg_world* w, uint32_t index, pxg_strand_set* out);
/* hair::set_strand_roots (hair/strand_builder.h:143-153) on set `index`: the first n strands'
 * roots move to positions (3 floats each) and, if given, orientations (x, y, z, w). */
int pxg_set_strand_roots(pxg_world* w, uint32_t index, const float* positions, const float* orientations, uint32_t n);
/* 32 x 32 layer matrix rows (broadphase/layer_matrix.h:21-47). */
int pxg_set_layer_masks(pxg_world* w, const uint32_t masks[32]);
/* Replaces the joint table: the joints_ array behind add_hinge_joint / add_ball_joint /
 * add_cone_twist_joint / add_joint / remove_joint / get_joint (world.h:400-518). Joints are
 * solved in table order (joint passes alternate direction per velocity iteration, solver.h:939-946);
 * body indices must be < the body count. n == 0 removes every joint. */
int pxg_set_joints(pxg_world* w, const pxg_joint* joints, uint32_t n);
/* Downloads the joint table with the solver's accumulators (get_joint, world.h:514). */
int pxg_get_joints(pxg_world* w, pxg_joint* out, uint32_t n);
/* Body pairs whose contacts are switched off regardless of layers: set_pair_collision(a, b,
 * false), world.h:527-532. ab holds n (a, b) body index pairs; n == 0 clears the set. */
int pxg_set_ignored_pairs(pxg_world* w, const uint32_t* ab, uint32_t n);

/* --- stepping --- */
int pxg_step_fixed(pxg_world* w, float h);                        /* step_fixed(), world.h:1278 */
/* Split step for the on_substep seam (world.h:1299): FORCES_COLLIDE runs gravity, force
 * integration, broadphase and narrowphase; SOLVE_INTEGRATE runs the solve, integration and sleep.
 * Bodies may be re-uploaded between the two. */
int pxg_step_phase(pxg_world* w, uint32_t phase, float h);
int pxg_step(pxg_world* w, float dt, uint32_t* substeps_run);     /* step(), world.h:1229 */
/* The once-per-frame cloth and hair pass that ends step() (world.h:1241-1256), for callers that
 * run the fixed steps themselves: frame is the clamped frame time, substeps_run the fixed steps
 * the frame ran (0: the broadphase boxes are synced over the frame first). pxg_step calls it. */
int pxg_step_soft(pxg_world* w, float frame, uint32_t substeps_run);
int pxg_synchronize(pxg_world* w);

/* --- results --- */
int pxg_get_stats(pxg_world* w, pxg_stats* out);
/* Broadphase pairs of the last step (shape slots, a < b), in discovery order. */
int pxg_get_pairs(pxg_world* w, uint32_t* ab, uint32_t cap, uint32_t* n);
/* Valid manifolds of the last step (manifolds(), world.h:731) in discovery order, with the
 * solver's final impulses. */
int pxg_get_manifolds(pxg_world* w, pxg_manifold* out, uint32_t cap, uint32_t* n);
/* The solver's order of the last step: manifold indices (into pxg_get_manifolds order), colour
 * by colour, followed by the non-trigger manifolds the solver skipped (no awake dynamic side).
 * colour_offsets[c] .. colour_offsets[c+1] delimit colour c; n_colours colours. */
int pxg_get_solve_order(pxg_world* w, uint32_t* order, uint32_t cap, uint32_t* n,
                        uint32_t* colour_offsets, uint32_t colour_cap, uint32_t* n_colours);
/* Per-shape fat AABBs (min xyz, max xyz) -- persistent broadphase state. */
int pxg_get_fat_aabbs(pxg_world* w, float* out6, uint32_t n);
int pxg_set_fat_aabbs(pxg_world* w, const float* in6, uint32_t n);
/* Warm-start cache (dynamics/solver.h:63-69): keys and (normal, tangent1, tangent2). */
int pxg_get_warm_cache(pxg_world* w, uint64_t* keys, float* impulses3, uint32_t cap, uint32_t* n);
int pxg_set_warm_cache(pxg_world* w, const uint64_t* keys, const float* impulses3, uint32_t n);
/* Contact events queued since the last clear (events(), world.h:744): every step_fixed appends
 * its events (emit_contact_events_, world.h:1998-2018: per cache -- solid, then trigger -- all
 * enters, then stays, then exits, each in pair order); pxg_step clears the queue before its
 * substeps (world.h:1230). n receives the queued count; min(cap, n) events are copied. Returns
 * PXG_ERR_CAPACITY (events still copied) when events were dropped because the queue was full. */
int pxg_get_events(pxg_world* w, pxg_contact_event* out, uint32_t cap, uint32_t* n);
int pxg_clear_events(pxg_world* w);                               /* clear_events(), world.h:748 */
/* Event generation on (default) or off, and the queue capacity in events (0 = keep; default
 * 2 * max_manifolds + 1024). Switching it on starts from an empty pair snapshot, so the next step
 * reports its contacts as enters. No counterpart in the reference, which always classifies. */
int pxg_set_events(pxg_world* w, uint32_t enabled, uint32_t capacity);
/* Scene raycasts against the bodies' current poses, culled by the broadphase's fat AABBs as the
 * reference's trees cull them (sphere, box and capsule shapes). pxg_raycast: the closest hit of
 * each of n rays (raycast(), world.h:766-791; equal distances go to the lower shape slot).
 * pxg_raycast_any: any[i] = 1 when ray i hits anything (raycast_any(), world.h:803-817).
 * pxg_raycast_all: every hit of one ray, nearest first (raycast_all(), world.h:825-847); n
 * receives the hit count, min(cap, n) hits are copied. Host buffers in and out; the queries run on
 * the world's stream after its pending steps. */
int pxg_raycast(pxg_world* w, const pxg_ray* rays, uint32_t n, const pxg_query_filter* filter, pxg_raycast_hit* hits);
int pxg_raycast_any(pxg_world* w, const pxg_ray* rays, uint32_t n, const pxg_query_filter* filter, uint32_t* any);
/* pxg_raycast for rays already in device memory (written before the call) and hits written to
 * device memory, enqueued on the world's stream without waiting: pxg_synchronize orders the
 * caller after it. For callers whose rays live on the GPU; no counterpart in the reference. */
int pxg_raycast_device(pxg_world* w, const pxg_ray* d_rays, uint32_t n, const pxg_query_filter* filter,
                       pxg_raycast_hit* d_hits);
int pxg_raycast_all(pxg_world* w, const pxg_ray* ray, const pxg_query_filter* filter, pxg_raycast_hit* hits,
                    uint32_t cap, uint32_t* n);
/* Every body with a shape overlapping the query shape (overlap_sphere / overlap_box /
 * overlap_capsule, world.h:1002-1060): the query's AABB against the fat AABBs (inclusive), the
 * filter, then the exact narrowphase test (query/queries.h:139-258, sphere, box and capsule
 * shapes). Bodies come once per overlapping shape, in shape-slot order; n receives the count,
 * min(cap, n) are copied. */
int pxg_overlap(pxg_world* w, const pxg_overlap_query* query, const pxg_query_filter* filter, uint32_t* bodies,
                uint32_t cap, uint32_t* n);
/* n overlap queries in one GPU pass (the batched form of pxg_overlap; the reference answers one
 * per call): counts[i] receives the number of overlapping shapes of queries[i], and
 * bodies[i * cap ... i * cap + min(cap, counts[i]) - 1] their owners, in shape-slot order, exactly
 * as pxg_overlap lists them. bodies holds n * cap entries. */
int pxg_overlap_batch(pxg_world* w, const pxg_overlap_query* queries, uint32_t n, const pxg_query_filter* filter,
                      uint32_t* bodies, uint32_t cap, uint32_t* counts);
/* pxg_overlap_batch for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). It has no counterpart in the
 * reference. d_queries (n records), d_bodies (n * cap entries) and d_counts (n entries) are
 * device pointers; the results are those of pxg_overlap_batch, except that a query whose type is
 * not a sphere, box or capsule gets the count PXG_NO_BODY. The call only enqueues work on the
 * world's stream and returns: the results are ready after pxg_synchronize (or any later wait on
 * the world's stream). It allocates nothing: its scratch is a fixed device array of the library,
 * one per device (about 4 MB), which limits it to worlds of at most 65536 shape slots
 * (PXG_ERR_CAPACITY beyond). Every call on one device shares that array: calls on one device that
 * overlap in time (two worlds' streams, two host threads) race and must be serialised by the
 * caller. */
int pxg_overlap_batch_noalloc(pxg_world* w, const pxg_overlap_query* d_queries, uint32_t n,
                              const pxg_query_filter* filter, uint32_t* d_bodies, uint32_t cap, uint32_t* d_counts);
/* The first body a swept sphere, box or capsule touches (shape_cast() / sphere_cast() /
 * box_cast() / capsule_cast(), world.h:877-950; query/sweep.h:422-505: conservative advancement on
 * the exact closest distance for round shapes, a swept separating-axis test for box against box,
 * a start overlap reported at distance 0 only when the sweep leads into the target). hit->hit = 0
 * when nothing is touched. Equal distances go to the lower shape slot. */
int pxg_shape_cast(pxg_world* w, const pxg_cast_query* cast, const pxg_query_filter* filter, pxg_raycast_hit* hit);
/* n shape casts in one GPU pass (the batched form of pxg_shape_cast, for callers that cast many
 * shapes per step, as vehicles and character motors do; the reference casts one per call):
 * hits[i] is exactly what pxg_shape_cast returns for casts[i]. */
int pxg_shape_cast_batch(pxg_world* w, const pxg_cast_query* casts, uint32_t n, const pxg_query_filter* filter,
                         pxg_raycast_hit* hits);
/* Every shape the posed query shape (sphere, box or capsule) penetrates, with the push-out that
 * separates them (compute_penetration(), world.h:963-995; query/sweep.h:339-384: the solver's
 * contact generation, with the capsule-against-box pair handled on the exact segment-box distance).
 * Candidates are live shapes whose fat AABB meets the query shape's bounds and that pass the filter;
 * only positive depths are reported, in shape-slot order. n receives the count, min(cap, n) are
 * copied. */
int pxg_compute_penetration(pxg_world* w, const pxg_shape* shape, const float position[3], const float orientation[4],
                            const pxg_query_filter* filter, pxg_penetration* out, uint32_t cap, uint32_t* n);
/* n compute_penetration queries in one GPU pass (the batched form of pxg_compute_penetration; the
 * reference answers one per call): probes[i] gives the shape, position and orientation (its
 * direction and max_distance are not used); counts[i] receives its number of penetrations and
 * out[i * cap ... i * cap + min(cap, counts[i]) - 1] the records, exactly as
 * pxg_compute_penetration returns them. out holds n * cap records. */
int pxg_compute_penetration_batch(pxg_world* w, const pxg_cast_query* probes, uint32_t n, const pxg_query_filter* filter,
                                  pxg_penetration* out, uint32_t cap, uint32_t* counts);
/* FNV-1a over alive bodies' position, orientation, velocities (world_state_hash, world.h:1334). */
int pxg_state_hash(pxg_world* w, uint64_t* out);

/* ---- learned warm start ----
 * A 24-64-64-3 ReLU network predicts the initial (normal, tangent1, tangent2) impulses of contact
 * points that have no cached impulse (the reference starts them at zero, dynamics/solver.h:166-170).
 * Predictions are clamped to lambda_n >= 0 and |lambda_t| <= mu lambda_n; the solver then runs as
 * usual, so the residual still judges convergence. Weights are FP32 here and BF16 on the device
 * (inference on the matrix cores with FP32 accumulation). Input i is (x_i - in_shift_i) *
 * in_scale_i; output k is y_k * out_scale_k. No counterpart in the reference.
 * Features, per contact point, from the pre-solve state:
 *   0-2 normal; 3-5 relative velocity (b - a) at the point in the contact frame (n, t1, t2);
 *   6 penetration; 7 separation; 8-9 effective inverse masses of a and b; 10-12 / 13-15 lever
 *   arms ra / rb in the contact frame; 16-18 / 19-21 Ia^-1 (ra x n) / Ib^-1 (rb x n) in the
 *   contact frame; 22 friction; 23 restitution. */
typedef struct pxg_mlp {
    float w1[64 * 24]; /* [out][in] */
    float b1[64];
    float w2[64 * 64];
    float b2[64];
    float w3[3 * 64];
    float b3[3];
    float in_shift[24];
    float in_scale[24];
    float out_scale[3];
} pxg_mlp;
typedef struct pxg_mlp_train_config {
    uint32_t epochs, batch, seed, reserved0;
    float lr, lr_decay, beta1, beta2, eps, weight_decay;
} pxg_mlp_train_config;
/* NULL turns the learned warm start off. */
int pxg_set_mlp(pxg_world* w, const pxg_mlp* net);
/* Records of the last step: features of every active contact point (24 floats), its final
 * impulses (3 floats) and whether it missed the warm-start cache. Requires pxg_set_recording(1)
 * before the step. */
int pxg_set_recording(pxg_world* w, uint32_t on);
int pxg_get_records(pxg_world* w, float* features, float* impulses, uint32_t* missed, uint32_t cap, uint32_t* n);
/* The learned predictions the last step applied: warm-start keys and clamped impulses. */
int pxg_get_predictions(pxg_world* w, uint64_t* keys, float* impulses3, uint32_t cap, uint32_t* n);
/* Standalone inference / gradient / training on the current device. */
int pxg_mlp_infer(const pxg_mlp* net, const float* features, uint32_t n, float* out3);
int pxg_mlp_gradients(const pxg_mlp* net, const float* X, const float* Y, uint32_t n, pxg_mlp* grad, float* loss);
int pxg_mlp_train(pxg_mlp* net, const float* X, const float* Y, uint32_t n, const pxg_mlp_train_config* cfg,
                  float* loss_per_epoch);
const char* pxg_mlp_last_error(void);

/* Profiling of the dominant kernel, the velocity pass: with profiling on, every velocity-pass
 * launch is bracketed by HIP events on the world's stream and counts its algorithmic bytes
 * (192 per manifold + 40 per contact point). pxg_get_profile returns and resets the totals. No
 * counterpart in the reference (it has no per-stage timers). */
typedef struct pxg_profile {
    double velocity_ms;          /* summed event time of velocity-pass launches */
    double velocity_bytes;       /* summed algorithmic bytes of those launches */
    uint64_t velocity_launches;
    /* PXG_ORDER_TILED solves, the workgroup of the first tile, summed over the launches, in 10 ns
     * ticks: running its colours, staging its bodies in and out of LDS, waiting at grid barriers,
     * computing and reducing residuals */
    uint64_t tiled_colour_ticks, tiled_stage_ticks, tiled_barrier_ticks, tiled_residual_ticks;
    uint64_t reserved[1];
} pxg_profile;
int pxg_set_profiling(pxg_world* w, uint32_t on);
int pxg_get_profile(pxg_world* w, pxg_profile* out);

const char* pxg_last_error(void);
const char* pxg_version(void);

/* ---- batched stepping of many independent scenes ----
 * Scenes are concatenated: scene k owns bodies body_offsets[k] .. body_offsets[k+1] and shapes
 * shape_offsets[k] .. shape_offsets[k+1]; a shape's owner is a scene-local body index. Each
 * scene behaves exactly like its own PhysicsWorld (world.h) stepped with step_fixed(h); one
 * pxg_batch_step call advances every scene `steps` times in one persistent launch. At most 288
 * bodies and 288 shapes per scene. No counterpart in the reference (it steps one world). */
typedef struct pxg_batch pxg_batch;
typedef struct pxg_batch_scene_stats {
    uint32_t pairs, manifolds, points, colours;
    uint32_t active_manifolds, velocity_passes, error, ordered;
    float residual;
    uint32_t awake_bodies;
    uint32_t learned_predictions;   /* warm-start misses the network predicted in the last step */
    uint32_t velocity_passes_total; /* velocity passes of every step since the batch was created */
} pxg_batch_scene_stats;
int pxg_batch_create(const pxg_config* cfg, uint32_t n_scenes, const uint32_t* body_offsets,
                     const uint32_t* shape_offsets, const pxg_body* bodies, const pxg_shape* shapes,
                     pxg_batch** out);
int pxg_batch_destroy(pxg_batch* b);
/* block_size 64/128/256/512 threads per scene; longest_first = take scenes by decreasing size */
int pxg_batch_set_tuning(pxg_batch* b, uint32_t block_size, uint32_t longest_first);
/* Workgroups (one scene each at a time, scenes from a queue) launched per compute unit: 0 = as
 * many as the kernel's LDS and registers allow; larger values are clamped to that. Tunings change
 * only the speed, never the results. */
int pxg_batch_set_occupancy(pxg_batch* b, uint32_t workgroups_per_cu);
int pxg_batch_set_layer_masks(pxg_batch* b, const uint32_t masks[32]);
int pxg_batch_step(pxg_batch* b, float h, uint32_t steps);
int pxg_batch_synchronize(pxg_batch* b);
int pxg_batch_get_bodies(pxg_batch* b, pxg_body* out, uint32_t n);
int pxg_batch_get_stats(pxg_batch* b, pxg_batch_scene_stats* out, uint32_t n);
int pxg_batch_get_manifolds(pxg_batch* b, uint32_t scene, pxg_manifold* out, uint32_t cap, uint32_t* n);
int pxg_batch_get_solve_order(pxg_batch* b, uint32_t scene, uint32_t* order, uint32_t cap, uint32_t* n,
                              uint32_t* colour_offsets, uint32_t colour_cap, uint32_t* n_colours);
/* Joint tables of every scene: scene k owns joints[joint_offsets[k] .. joint_offsets[k+1]) with
 * scene-local body indices (as pxg_set_joints for one world, world.h:400-518). */
int pxg_batch_set_joints(pxg_batch* b, const uint32_t* joint_offsets, const pxg_joint* joints);
int pxg_batch_get_joints(pxg_batch* b, pxg_joint* out, uint32_t n);
/* Ignored body pairs of every scene (set_pair_collision(a, b, false), world.h:527-532): scene k
 * owns the pairs ab[2 pair_offsets[k] .. 2 pair_offsets[k+1]), scene-local. */
int pxg_batch_set_ignored_pairs(pxg_batch* b, const uint32_t* pair_offsets, const uint32_t* ab);
/* The learned warm start in every scene (as pxg_set_mlp for one world); NULL turns it off. At most
 * 256 misses per scene and step are predicted, the rest start at zero as in the reference. */
int pxg_batch_set_mlp(pxg_batch* b, const pxg_mlp* net);
/* The last step's predictions of one scene, sorted by warm-start key (scene-local bodies). */
int pxg_batch_get_predictions(pxg_batch* b, uint32_t scene, uint64_t* keys, float* impulses3, uint32_t cap,
                              uint32_t* n);
const char* pxg_batch_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* PXG_H */
