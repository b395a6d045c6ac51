/*
 * physecs_gpu.h: C ABI of the MI355X (gfx950) rigid-body step.
 *
 * This is the drop-in boundary under physecs::Scene::simulate(float): host code gathers
 * the EnTT components (TransformComponent, RigidBodyDynamicComponent,
 * RigidBodyCollisionComponent) into the scene descriptor below, creates a world, and
 * steps it; the world keeps all state resident in HBM between steps.
 *
 * Conventions: every function returns an int status (PGX_OK = 0, negative on error;
 * pgx_last_error() describes the last error of the calling thread). No C++ exception
 * crosses this boundary. Host pointers are borrowed for the duration of the call only.
 * Device memory is owned by the world handle. One HIP stream per world; a world is not
 * thread-safe; distinct worlds are independent.
 *
 * Reference interfaces replaced (paths relative to the reference repository):
 *   pgx_world_create        Scene::Scene + the registry signal hooks that register
 *                           bodies (include/Physecs/Physecs.h:199, src/Physecs.cpp:25-37,92-98)
 *   pgx_world_set_params    Scene::setNumSubSteps / setNumIterations / setGravity
 *                           (include/Physecs/Physecs.h:201-203)
 *   pgx_world_step          Scene::simulate(float timeStep) (include/Physecs/Physecs.h:204,
 *                           src/Physecs.cpp:112-561)
 *   pgx_world_set_can_collide  Scene::setCanCollide (include/Physecs/Physecs.h:223)
 *   pgx_world_add_joints    Scene::createJoint<T> for the built-in joint types
 *                           (include/Physecs/Physecs.h:209-214, src/Physecs.cpp:690-710)
 *   pgx_world_set_joint_params  the joints' setters (include/Physecs/Joints/<name>Joint.h)
 *   pgx_world_set_joint_callbacks  user Joint subclasses (include/Physecs/Joint.h:42-70):
 *                           getSolverDesc per step, MakeConstraintsFunc per substep
 *                           (src/Physecs.cpp:321-355,430-439)
 *   pgx_world_destroy_joint Scene::destroyJoint (include/Physecs/Physecs.h:215,
 *                           src/Physecs.cpp:712-725)
 *   pgx_world_read_state / pgx_world_write_state
 *                           the in-place component writes that simulate performs
 *                           (Transform position/orientation, Dynamic velocity/angularVelocity)
 *   pgx_server_*            no single reference counterpart: many independent Scenes, each
 *                           simulate()d once per server step (include/Physecs/Physecs.h:199-204),
 *                           batched into one device world with continuous admission
 */
#ifndef PHYSECS_GPU_H
#define PHYSECS_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PGX_OK 0
#define PGX_ERR_INVALID -1      /* bad argument */
#define PGX_ERR_HIP -2          /* HIP runtime error (no device, out of memory, launch) */
#define PGX_ERR_CAPACITY -3     /* a per-step buffer overflowed; raise the capacity */
#define PGX_ERR_UNSUPPORTED -4  /* shape or feature not supported on the device path */

/* Geometry types, include/Physecs/Colliders.h:9 */
#define PGX_SPHERE 0
#define PGX_CAPSULE 1
#define PGX_BOX 2
#define PGX_CONVEX_MESH 3
#define PGX_TRIANGLE_MESH 4

typedef struct pgx_world pgx_world;

typedef struct pgx_config {
    int num_substeps;       /* Scene::numSubSteps, default 8 (Physecs.h:156) */
    int num_iterations;     /* Scene::numIterations, default 2 (Physecs.h:157) */
    float gravity;          /* Scene::g, default 9.81 along -y (Physecs.h:159) */
    int device;             /* HIP device ordinal */
    int pair_capacity;      /* candidate pairs per step; 0 = 12 x colliders */
    int block_size;         /* threads per contact-solve workgroup at creation: 64, 128 or 256;
                               0 = 64 (pgx_tuning.solve_block; the autotuner may change it) */
} pgx_config;

/* Scene descriptor. Actors are entities with a TransformComponent; the dynamic storage
 * holds RigidBodyDynamicComponent entries in storage order (its index is the body index
 * `b` the reference's solver uses); colliders are listed in broadphase-entry creation
 * order. Arrays are tightly packed float32 / int32. Quaternions are x,y,z,w. Inverse
 * inertia tensors are 3x3 column-major (glm::mat3). Material = friction, restitution,
 * damping. col_params: sphere {r}, capsule {halfHeight, r}, box {hx, hy, hz}.
 * actor_world may be NULL (one world); otherwise worlds never interact. */
typedef struct pgx_scene_desc {
    int n_actors;
    const uint32_t* entity;       /* entt::entity value, n_actors */
    const float* pos;             /* 3 x n_actors */
    const float* quat;            /* 4 x n_actors */
    const int32_t* dyn_of_actor;  /* dynamic-storage index or -1 */
    int n_dyn;
    const int32_t* actor_of_dyn;
    const uint8_t* kinematic;
    const float* vel;             /* 3 x n_dyn */
    const float* angvel;          /* 3 x n_dyn */
    const float* inv_mass;        /* n_dyn */
    const float* com;             /* 3 x n_dyn, body-local centre of mass */
    const float* inv_inertia;     /* 9 x n_dyn, body-local */
    int n_colliders;
    const int32_t* col_actor;
    const int32_t* col_index;     /* index within the actor's collider list */
    const float* col_pos;         /* 3 x n_colliders, actor-local */
    const float* col_quat;        /* 4 x n_colliders */
    const int32_t* col_type;
    const float* col_params;      /* 4 x n_colliders */
    const float* col_material;    /* 3 x n_colliders */
    const uint8_t* col_trigger;
    const uint8_t* col_enable;    /* Collider::enableSimulation */
    const int32_t* col_data;
    int n_worlds;
    const int32_t* actor_world;   /* world of each actor, or NULL */
    /* Convex meshes (include/Physecs/ConvexMesh.h:7-34). A CONVEX_MESH collider's params
     * are {scale.x, scale.y, scale.z, mesh index}. Mesh m owns vertices
     * [convex_vert_start[m], convex_vert_start[m+1]) and faces
     * [convex_face_start[m], convex_face_start[m+1]); face f lists the mesh-local vertex
     * indices face_idx[face_idx_start[f] .. face_idx_start[f+1]), clockwise seen from
     * outside (the winding the reference's clipping keeps, src/Clipping.cpp:9-11), with
     * its outward unit normal and centroid. */
    int n_convex;
    const int32_t* convex_vert_start;  /* n_convex + 1 */
    const float* convex_verts;         /* 3 x vertices */
    const int32_t* convex_face_start;  /* n_convex + 1 */
    const int32_t* face_idx_start;     /* faces + 1 */
    const int32_t* face_idx;
    const float* face_normal;          /* 3 x faces */
    const float* face_centroid;        /* 3 x faces */
    /* Static triangle meshes (include/Physecs/TriangleMesh.h). A TRIANGLE_MESH collider's
     * params are {0, 0, 0, mesh index}; mesh m owns vertices [tri_vert_start[m],
     * tri_vert_start[m+1]) and triangles [tri_start[m], tri_start[m+1]) given as three
     * mesh-local vertex indices each, in the order TriangleMesh's constructor takes them. */
    int n_trimesh;
    const int32_t* tri_vert_start;     /* n_trimesh + 1 */
    const float* tri_verts;            /* 3 x vertices */
    const int32_t* tri_start;          /* n_trimesh + 1 */
    const int32_t* tri_idx;            /* 3 x triangles */
    /* Optional (NULL, or NULL entries: build at creation): per triangle mesh, the bytes
     * pgx_trimesh_cache wrote for that same mesh, and their lengths (required with
     * tri_bvh_cache); creation then loads its triangle order and BVH instead of building
     * them. A cache of another mesh, of another length, or with a corrupted payload is
     * rejected (pgx_trimesh_cache_validate). */
    const void* const* tri_bvh_cache;
    const size_t* tri_bvh_cache_bytes;
} pgx_scene_desc;

/* Built-in joint types (include/Physecs/Joints/<name>Joint.h) */
#define PGX_JOINT_SPHERICAL 0
#define PGX_JOINT_UNIVERSAL 1
#define PGX_JOINT_REVOLUTE 2
#define PGX_JOINT_FIXED 3
#define PGX_JOINT_PRISMATIC 4
#define PGX_JOINT_GEAR 5
#define PGX_JOINT_SERVO 6
/* A user Joint subclass (include/Physecs/Joint.h:42-70): its rows come from the world's
 * joint callbacks (pgx_world_set_joint_callbacks); only type, actors and anchors are read. */
#define PGX_JOINT_CUSTOM 7

/* Constraint1D flags (include/Physecs/Constraint1D.h:11-16). A row has one of NONE,
 * ANGULAR, SOFT, LIMITED, ANGULAR | SOFT, ANGULAR | LIMITED: the combinations the
 * reference's constraint containers hold (include/Physecs/Constraint1DContainer.h:117-118). */
#define PGX_ROW_NONE 0
#define PGX_ROW_SOFT 1
#define PGX_ROW_ANGULAR 2
#define PGX_ROW_LIMITED 4
#define PGX_JOINT_MAX_ROWS 7    /* rows of one joint on the device */

/* JointWorldSpaceData (include/Physecs/Joint.h:16-23), the same memory layout: the anchors
 * in world space (p0, p1), the anchors relative to the bodies' centres of mass rotated into
 * world space (r0, r1), and the anchor frames (u0, u1; 3x3, column-major as glm::mat3),
 * computed as JointSolverData::calculateWorldSpaceData does (src/Joint.cpp:5-14). */
typedef struct pgx_joint_world_space {
    float p0[3], p1[3], r0[3], r1[3];
    float u0[9], u1[9];
} pgx_joint_world_space;

/* The fields of one Constraint1D that MakeConstraintsFunc writes through Constraint1DView
 * (include/Physecs/Constraint1DContainer.h:22-91). */
typedef struct pgx_joint_row {
    float linear[3], angular0[3], angular1[3];
    float target_velocity, c, min, max, frequency, damping_ratio;
} pgx_joint_row;

/* Joint::getSolverDesc, for joint `joint` (creation order) at the start of each step
 * (src/Physecs.cpp:321-355): writes its rows' flags (PGX_ROW_*, in creation order) into
 * row_flags and returns their number, 0 to PGX_JOINT_MAX_ROWS; a negative return fails
 * the step. `ws` holds the joint's world-space data at the step's starting poses. */
typedef int (*pgx_joint_layout_fn)(void* user, int joint, const pgx_joint_world_space* ws, int32_t* row_flags);
/* MakeConstraintsFunc, for joint `joint` at every substep (src/Physecs.cpp:430-439): `rows`
 * holds the joint's n_rows rows in creation order, as the previous call of this step left
 * them (Constraint1D's defaults at the step's first substep: zeros, min = -FLT_MAX,
 * max = FLT_MAX; include/Physecs/Constraint1D.h:19-35); the function sets what it writes. */
typedef void (*pgx_joint_rows_fn)(void* user, int joint, const pgx_joint_world_space* ws, pgx_joint_row* rows,
                                  int n_rows);

/* One Scene::createJoint<T>(entity0, anchor0Pos, anchor0Or, entity1, anchor1Pos, anchor1Or).
 * Anchors are in body space: position xyz, orientation xyzw. The other fields are the
 * joint's data members, as its setters set them; a field a type does not have is ignored.
 * Reference defaults: PrismaticJoint.h:8-17 (upper 1, lower 0, stiffness 5, damping 1),
 * ServoJoint.h:7-11 (target 0, stiffness 30, damping 1), GearJoint.h:7-14 (ratio 1),
 * RevoluteJoint.h:8-12 (drive off, 0, 0). */
typedef struct pgx_joint_desc {
    int32_t type;
    int32_t actor0, actor1;   /* descriptor actor indices */
    float anchor0[7];
    float anchor1[7];
    int32_t drive_enabled;    /* revolute, prismatic: setDriveEnabled */
    float drive_velocity;     /* revolute: setDriveVelocity */
    float drive_max_torque;   /* revolute: setDriveMaxTorque */
    float lower_limit;        /* prismatic: setLowerLimit */
    float upper_limit;        /* prismatic: setUpperLimit */
    float target;             /* prismatic: setTargetPosition; servo: setTargetAngle */
    float drive_stiffness;    /* prismatic, servo: setDriveStiffness */
    float drive_damping;      /* prismatic, servo: setDriveDamping */
    float gear_ratio;         /* gear: setGearRatio */
} pgx_joint_desc;

typedef struct pgx_counts {
    int pairs;        /* broadphase candidate pairs of the last step */
    int manifolds;    /* contact manifolds with at least one point */
    int points;       /* contact points */
    int colors;       /* graph colors used by the contact solve */
    int joint_rows;   /* joint constraint rows */
    int coloring_rounds;
} pgx_counts;

/* Library / device */
const char* pgx_last_error(void);
int pgx_device_count(int* count);

/* Diagnostic, no counterpart in the reference: the library's stable radix sort of (keys,
 * vals) by the low `bits` bits of the keys, the sort the step uses for Morton codes, pair
 * keys, shape-pair classes and colours. Host arrays in and out; runs on `device`. */
int pgx_sort_u64(int device, const uint64_t* keys, const int32_t* vals, int n, int bits, uint64_t* keys_out,
                 int32_t* vals_out);
/* No counterpart in the reference. For callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops that must not call hipMalloc): pgx_sort_u64 on
 * device arrays, sorted in place, asynchronously on `stream` (a hipStream_t; NULL for the
 * null stream), allocating nothing. The sort's scratch is a static 12.3 MiB array each GPU
 * holds until the process ends, so n is at most 1048576 (PGX_ERR_INVALID beyond), and calls
 * that overlap on one GPU race for it. */
int pgx_sort_u64_noalloc(int device, uint64_t* keys, int32_t* vals, int n, int bits, void* stream);

/* World lifetime. The descriptor is copied to the device; creation-time bounds carry no
 * margin (src/Physecs.cpp:32). cfg may be NULL for the defaults. */
int pgx_world_create(const pgx_config* cfg, const pgx_scene_desc* scene, pgx_world** out);
int pgx_world_destroy(pgx_world* w);
int pgx_world_set_params(pgx_world* w, int num_substeps, int num_iterations, float gravity);
/* Scene::setContactFilter (include/Physecs/Physecs.h:174,224; src/Physecs.cpp:200,795-796)
 * for a filter function ContactType(*)(bool isTrigger0, int data0, bool isTrigger1, int
 * data1), tabulated by the caller: collider_class[c] (descriptor order) in [0, n_classes)
 * names each collider's (isTrigger, data) value, and table[k0 * n_classes + k1] != 0 is
 * the filter returning TRIGGER for a pair whose lower entity's collider has class k0 and
 * whose other collider has class k1 (the argument order of src/Physecs.cpp:200). TRIGGER
 * pairs get the overlap test and enter/exit events instead of contacts. n_classes 0 or
 * table NULL restores defaultContactFilter (a trigger on either side). Takes effect from
 * the next step. A table that makes trigger pairs gives the world trigger lists, which
 * its saved states then carry. */
#define PGX_MAX_FILTER_CLASSES 1024
int pgx_world_set_contact_filter(pgx_world* w, int n_classes, const uint8_t* table, const int32_t* collider_class);
int pgx_world_set_can_collide(pgx_world* w, uint32_t entity0, uint32_t entity1, int can_collide);
/* Appends joints in creation order. Each takes the lowest joint colour free on both of its
 * entities (8 colours, then the overflow colour), and the pair stops colliding. */
int pgx_world_add_joints(pgx_world* w, int n, const pgx_joint_desc* joints);
/* Sets the data members of joint `index` (creation order) from `joint`'s type-specific
 * fields, as the joint's setters do; type, actors and anchors must be unchanged. Takes
 * effect from the next step. */
int pgx_world_set_joint_params(pgx_world* w, int index, const pgx_joint_desc* joint);
/* Scene::destroyJoint: the joint leaves its colour, its colour bit is freed on both
 * entities and the entity pair collides again (even where another joint still joins
 * them, as the reference does). Other joints keep their indices. */
int pgx_world_destroy_joint(pgx_world* w, int index);
/* The host functions of the world's PGX_JOINT_CUSTOM joints (user Joint subclasses,
 * include/Physecs/Joint.h:42-70; the calls of src/Physecs.cpp:321-355,430-439). Each step
 * calls `layout` for every live custom joint, joint colour by colour and in creation order
 * inside a colour, as the reference walks its joint colours; each substep calls `rows` in the
 * same order. The device waits for these calls: the step stops at its start and at each
 * substep to hand the joints' world-space data to the host and take their rows back, so a
 * world with custom joints does not run its substep loop as a HIP graph. A step of a world
 * with custom joints and no callbacks fails with PGX_ERR_INVALID. NULL functions remove the
 * callbacks. The functions run inside pgx_world_step, on the calling thread, and must not
 * call into the same world. */
int pgx_world_set_joint_callbacks(pgx_world* w, pgx_joint_layout_fn layout, pgx_joint_rows_fn rows, void* user);

/* One Scene::simulate(time_step) on the device. Asynchronous; pgx_world_sync waits. */
int pgx_world_step(pgx_world* w, float time_step);
int pgx_world_sync(pgx_world* w);

/* Host <-> device body state, in the descriptor's actor and dynamic-storage order.
 * pos: 3 x n_actors, quat: 4 x n_actors, vel/angvel: 3 x n_dyn. Any pointer may be NULL. */
int pgx_world_read_state(pgx_world* w, float* pos, float* quat, float* vel, float* angvel);
int pgx_world_write_state(pgx_world* w, const float* pos, const float* quat, const float* vel, const float* angvel);

/* Introspection of the last step (for tests and tools). */
int pgx_world_get_counts(pgx_world* w, pgx_counts* out);
/* Bounds of every collider, descriptor order: min xyz, max xyz. */
int pgx_world_get_bounds(pgx_world* w, float* out);
/* Candidate pairs as descriptor collider indices (collider of the smaller entity first),
 * in the device's canonical order. Writes at most cap pairs; returns count in *n. */
int pgx_world_get_pairs(pgx_world* w, int32_t* out, int cap, int* n);
/* Manifolds in solve order (graph color order). keys: e0, c0, e1, c1, triangleIndex;
 * ints: b0, b1, numPoints, color; floats (56): normal(3), friction, 4 x (p0, p1) world
 * points, 4 x (r0 local, r1 local, targetVelocity). Any pointer may be NULL. */
int pgx_world_get_manifolds(pgx_world* w, int32_t* keys, int32_t* ints, float* floats, int cap, int* n);

/* Timing of the phases of the last timed step on the device, in microseconds (HIP events
 * on the world's stream). pgx_world_enable_timing(w, n): 0 off, 1 every step, n > 1 every
 * n-th step from the next one on (the steps between record no events and run as without
 * timing: a timed step is slower, by its per-launch events and its wait at the end). */
int pgx_world_enable_timing(pgx_world* w, int enable);
int pgx_world_get_timing(pgx_world* w, float* broadphase_us, float* narrowphase_us, float* coloring_us,
                         float* solver_us, float* total_us);

/* Cumulative timing since pgx_world_reset_stats, collected on the timed steps.
 * Default (per-colour) mode: every k_solve launch of the first substep of each step is
 * bracketed by HIP events on the world's stream; solve_bytes is their algorithmic traffic,
 * 284 B per contact point + 12 B per manifold of the launched colour (rows 108 + scalars
 * 24 + solver and component velocities of both bodies 96 read; lambdas 8 + velocities 48
 * written). Persistent mode: the substep-loop kernel (one cooperative launch per step) is
 * bracketed instead, and solve_bytes follows the per-substep phase model (update rows
 * 164 B/manifold + 96 B/point, integrate velocities 168 B/body, preSolve 80 B/manifold +
 * 116 B/point, each solve pass 284 B/point + 12 B/manifold, integrate positions 120 B/body,
 * write-back 48 B/body; per substep). */
typedef struct pgx_stats {
    long steps;
    long solve_launches;
    double solve_us;
    double solve_bytes;
    double step_us;
    /* the runs of small colours solved in one workgroup (k_solve_batch), bracketed and
     * counted the same way over their colours' points and manifolds */
    long batch_launches;
    double batch_us;
    double batch_bytes;
} pgx_stats;
int pgx_world_reset_stats(pgx_world* w);
/* Execution of the substep loop's contact passes (the results are the same in every mode):
 *   PGX_EXEC_COLORS (default) one launch per colour and pass
 *   PGX_EXEC_FLOW            one launch per pass; each manifold solves as soon as its
 *                            bodies have taken every lower colour's solves
 *   PGX_EXEC_FLOW_XCD        the same with each world's manifolds on one XCD
 *   PGX_EXEC_COOPERATIVE     every substep of a step in one cooperative launch, grid
 *                            barriers between colours (worlds without joints; others use
 *                            PGX_EXEC_COLORS). blocks = its grid size (0 = one per CU).
 *   PGX_EXEC_WORLDS          one launch per pass, one workgroup per world: the world's
 *                            solver velocities in LDS, its colours in order with
 *                            workgroup barriers. Needs each world's dynamic bodies to be
 *                            contiguous and to fit in LDS (5120 bodies on gfx950);
 *                            otherwise the step uses PGX_EXEC_COLORS.
 *   PGX_EXEC_PASS            one cooperative launch per pass, every colour in it with a
 *                            grid barrier between colours; the solver velocities go
 *                            through the device's coherence point, so a barrier needs no
 *                            cache writeback. blocks = its grid size (0 = enough for the
 *                            largest colour, at most one workgroup per CU).
 *   PGX_EXEC_XPASS           one cooperative launch per pass on the CUs of one XCD (the
 *                            one most of the launch's workgroups landed on), every colour
 *                            in it with a barrier between colours; the solver velocities
 *                            stay in that XCD's L2. blocks = its grid size over all XCDs
 *                            (0 = 8 x enough for the largest colour; at most 248). */
#define PGX_EXEC_COLORS 0
#define PGX_EXEC_COOPERATIVE 1
#define PGX_EXEC_FLOW 2
#define PGX_EXEC_FLOW_XCD 3
#define PGX_EXEC_WORLDS 4
#define PGX_EXEC_PASS 5
#define PGX_EXEC_XPASS 6
int pgx_world_set_exec(pgx_world* w, int mode, int blocks);
int pgx_world_get_exec(pgx_world* w, int* mode);
int pgx_world_get_stats(pgx_world* w, pgx_stats* out);

/* Scene::raycastClosest (include/Physecs/Physecs.h:216-217, src/Physecs.cpp:571-621) for n
 * rays at once, between steps, on the world's current poses and broadphase bounds. Ray i
 * starts at origins[3i..], runs along dirs[3i..] (not normalised: distances are in units
 * of |dir|, as in the reference) up to max_distance[i], in world ray_world[i] (NULL: world
 * 0), and rejects the entity ignore_entity[i] (NULL or 0xffffffff: none), the form of the
 * reference's filter argument this boundary takes; the reference's unfiltered overload
 * passes an empty filter, which its leaf test never accepts, so it reports no hit.
 * Results: hit_collider[i] = the closest hit collider (descriptor index) or -1, and
 * hit_distance[i] its distance t (the reference's hitPos = origin + dir * t). Equal
 * distances go to the lower collider index (the reference: its tree's shape). Triangle
 * meshes are never hit (src/Raycast.cpp:148). device_us (may be NULL): the query's device
 * time, tree build included, by HIP events. */
int pgx_world_raycast(pgx_world* w, int n, const float* origins, const float* dirs, const float* max_distance,
                      const int32_t* ray_world, const uint32_t* ignore_entity, int32_t* hit_collider,
                      float* hit_distance, double* device_us);
/* Scene::overlap (include/Physecs/Physecs.h:207, src/Physecs.cpp:628-651) for n query
 * shapes at once, between steps, on the world's current poses and broadphase bounds. Query
 * i is a shape of types[i] (PGX_SPHERE .. PGX_TRIANGLE_MESH) with params[4i..] as in
 * pgx_scene_desc.col_params (a convex mesh: scale xyz, then the index of one of the scene's
 * convex meshes), at positions[3i..] with orientations[4i..] (x, y, z, w), in world
 * query_world[i] (NULL: world 0). filters[i] != 0 keeps the colliders whose
 * pgx_scene_desc.col_data has a bit in common with it (NULL or 0: all). A collider is hit
 * when the query's bounds meet its broadphase bounds (inclusive) and physecs::overlap
 * (src/Overlap.cpp:203-235) says the shapes overlap; triangle meshes on either side never
 * do, and the box-convex test places the convex support at the box's position, as the
 * reference's does (src/Overlap.cpp:200). Results: query i's hits are
 * colliders[offsets[i] .. offsets[i+1]) (descriptor indices, ascending; the reference's
 * OverlapHit {entity, colIndex} follows from col_actor and the collider's index within its
 * actor; the reference lists them in its tree's order), offsets has n + 1 entries and
 * *total = offsets[n]. When *total exceeds capacity, offsets and *total are set, colliders
 * is not written and PGX_ERR_CAPACITY is returned: call again with capacity >= *total.
 * device_us (may be NULL): the query's device time, tree build and sort included. */
int pgx_world_overlap(pgx_world* w, int n, const int32_t* types, const float* params, const float* positions,
                      const float* orientations, const int32_t* query_world, const int32_t* filters, int capacity,
                      int32_t* offsets, int32_t* colliders, int32_t* total, double* device_us);
/* Scene::overlapWithMinTranslationalDistance (include/Physecs/Physecs.h:208,
 * src/Physecs.cpp:653-688) for n query shapes at once, between steps. Queries as in
 * pgx_world_overlap (sphere, capsule, box or convex mesh; a triangle-mesh query shape is
 * PGX_ERR_UNSUPPORTED), no filter. Every collider whose broadphase bounds meet the query's
 * bounds is tested with physecs::collision(collider, query); each manifold with points is a
 * hit: the collider, the manifold's normal (from the collider toward the query) and its mtd,
 * the largest dot(contact0 - contact1, normal) over its points and at least 0 (the
 * reference's OverlapMtdHit {entity, colIndex, normal, mtd}). Query i's hits are
 * [offsets[i], offsets[i+1]): colliders[k], normal_mtd[4k..4k+3] = normal xyz, mtd;
 * ordered by collider index, then in collision()'s manifold order (the reference: its
 * tree's order). When *total exceeds capacity, only *total is set and PGX_ERR_CAPACITY is
 * returned. device_us (may be NULL): the query's device time. */
int pgx_world_overlap_mtd(pgx_world* w, int n, const int32_t* types, const float* params, const float* positions,
                          const float* orientations, const int32_t* query_world, int capacity, int32_t* offsets,
                          int32_t* colliders, float* normal_mtd, int32_t* total, double* device_us);
/* One character: the CharacterController's members (include/Physecs/CharacterController.h:
 * 17-24) and its entity's pose. */
typedef struct pgx_character {
    float position[3];        /* the entity's TransformComponent::position, updated by a move */
    float orientation[4];     /* its orientation (x, y, z, w) */
    float height, radius;
    float max_step_height;    /* kept for the reference's signature; see below */
    float slope_limit;        /* cosine, the reference's default cos(45 degrees) */
    int32_t num_iterations;   /* the reference's default 2 */
    int32_t world;            /* the world the character moves in */
} pgx_character;
/* registry.patch<TransformComponent> of n entities between steps: the Scene's on_update
 * hook (src/Physecs.cpp:51-54,95) recomputes the bounds of every collider of a moved entity,
 * static or dynamic, with the 0.01 margin (Scene::updateBounds, :79-90). actors[i] (descriptor
 * order) takes positions pos[3i..] and orientations quat[4i..] (x, y, z, w). Velocities are
 * unchanged. (pgx_world_write_state writes poses as the step does, in place, without a hook.) */
int pgx_world_patch_transforms(pgx_world* w, int n, const int32_t* actors, const float* pos, const float* quat);
/* CharacterController::move (src/CharacterController.cpp:5-45) for n characters at once,
 * between steps: position += translations[3i..], then num_iterations rounds of
 * overlapWithMinTranslationalDistance with the character's capsule (half height
 * height / 2, radius, centred height / 2 + radius above the position along its up axis),
 * each hit pushing the position out: along up by mtd / (n . up) where n . up exceeds
 * slope_limit, else along n by mtd, in hit order (collider index, then manifold order;
 * the reference: its tree's order). flags[i] = CharacterCollisionFlags (DOWN 1, SIDES 2,
 * UP 4). The reference's step-up branch (:13-23) casts its rays through the unfiltered
 * raycastClosest, which never reports a hit, so it never changes the translation:
 * max_step_height and delta_time have no effect, as there. */
int pgx_world_character_move(pgx_world* w, int n, pgx_character* characters, const float* translations,
                             float delta_time, uint8_t* flags);
/* The last step's trigger events: Scene's onTriggerEnter / onTriggerExit listeners
 * (include/Physecs/Physecs.h:30-40,220-223; src/Physecs.cpp:199-206,538-552). A pair with a
 * trigger collider on either side (defaultContactFilter) makes no contact; when
 * physecs::overlap says its shapes overlap it is in the step's trigger set. enter = pairs
 * in this step's set but not the last step's, exit = the reverse. Each event is a pair
 * (collider of the lower entity, collider of the other), descriptor indices, as the
 * reference's ContactPair orders it; pairs ascend (the reference calls its listeners in
 * unordered_set order). enter and exit hold 2 * capacity ints; when either list exceeds
 * capacity, the counts are set and PGX_ERR_CAPACITY is returned. Events are of the last
 * pgx_world_step only: a caller that wants every event reads them after every step. */
int pgx_world_get_trigger_events(pgx_world* w, int capacity, int32_t* enter, int32_t* n_enter, int32_t* exit,
                                 int32_t* n_exit);
/* Host-side asset pipeline. physecs::setMassProps (include/Physecs/MassUtil.h:7-12,
 * src/MassUtil.cpp:74-192) for one body of n colliders: types, params (4 each, as
 * pgx_scene_desc.col_params), local poses (3 + 4 each), is_trigger (NULL: none; triggers
 * add no mass). Convex colliders' meshes come from the tables given (as in
 * pgx_scene_desc; NULL when there are none). Out: inv_mass, com[3], inv_inertia[9]
 * (column-major), bit for bit as the reference computes them, quirks included: a box's
 * volume is the product of its half extents, convex colliders' orientations are ignored,
 * and the fan centre of a convex mesh sums its vertex buffer padded to a multiple of 4.
 * Runs on the host; needs no device. */
int pgx_mass_properties(int n, const int32_t* types, const float* params, const float* local_pos,
                        const float* local_quat, const uint8_t* is_trigger, int n_convex,
                        const int32_t* convex_vert_start, const float* convex_verts,
                        const int32_t* convex_face_start, const int32_t* face_idx_start, const int32_t* face_idx,
                        float mass, float* inv_mass, float* com, float* inv_inertia);
/* Host-side asset pipeline: ConvexMesh data (include/Physecs/ConvexMesh.h:7-34) for the
 * convex hull of n >= 4 points (3 floats each; the reference has no hull builder and takes
 * vertices and faces as given). Out: the hull's vertices (those input points, ascending
 * input order; 3 * n floats of room), faces as polygons of coplanar hull triangles,
 * clockwise seen from outside: face f is face_idx[face_start[f] .. face_start[f+1]) (room:
 * 2n + 1 and 6n ints), outward unit normals and vertex-mean centroids (3 floats per face,
 * room for 2n faces). PGX_ERR_INVALID for fewer than 4 points or a flat cloud. Host code. */
int pgx_convex_hull(const float* points, int n, float* verts, int* n_verts, int* face_start, int* face_idx,
                    float* normals, float* centroids, int* n_faces);
/* Host-side asset pipeline: the TriangleMesh constructor's build (src/TriangleMesh.cpp:7-164:
 * triangle order, normals, centroids, SAH BVH) of one mesh as bytes, for
 * pgx_scene_desc.tri_bvh_cache, so worlds with this mesh skip the build at creation.
 * *bytes = the cache's size; with out NULL only the size is computed; PGX_ERR_CAPACITY when
 * cap < *bytes. The bytes hold a hash of the vertices and indices they were built from. */
int pgx_trimesh_cache(const float* verts, int n_verts, const int32_t* tris, int n_tris, void* out, size_t cap,
                      size_t* bytes);
/* Host code, no counterpart in the reference: checks `bytes` bytes of a pgx_trimesh_cache
 * cache against the mesh it is meant for, as world creation does before the device sees
 * it: length, the hashes of the mesh and of the payload, child and leaf ranges, vertex
 * indices, depth. PGX_OK or PGX_ERR_INVALID. */
int pgx_trimesh_cache_validate(const float* verts, int n_verts, const int32_t* tris, int n_tris, const void* cache,
                               size_t bytes);
/* Host code: the cache of a TriangleMesh that is already built (src/TriangleMesh.cpp:7-164
 * left its triangles in build order with normals and centroids, and its BVH nodes as
 * {bounds min, max, triCount, index}): the drop-in's path for the app's TriangleMesh
 * objects, whose input order is gone. `tris` (build order) are the descriptor's tri_idx for
 * this mesh. The same size and capacity rules as pgx_trimesh_cache; PGX_ERR_INVALID for a
 * node list that does not describe a tree over the triangles. */
int pgx_trimesh_cache_pack(const float* verts, int n_verts, const int32_t* tris, int n_tris, const float* tri_normals,
                           const float* tri_centroids, const float* node_bounds, const int32_t* node_tri_count,
                           const int32_t* node_index, int n_nodes, void* out, size_t cap, size_t* bytes);
/* Save and resume a world's state between steps (no counterpart in the reference, which
 * keeps its hidden state -- contactCache, GearJointData, triggerCache; include/Physecs/
 * Physecs.h:167-171, include/Physecs/Joints/GearJoint.h:7-14 -- only inside a live Scene).
 * The state is one byte string: poses, velocities, broadphase bounds, the contact cache, the
 * gear joints' angles, the trigger lists and the last step's sizes. pgx_world_state_size
 * gives its length now (it changes from step to step); pgx_world_save_state writes it
 * (PGX_ERR_CAPACITY when bytes is too small); pgx_world_load_state replaces the world's
 * state with one saved by a world created from the same descriptor and joints
 * (PGX_ERR_INVALID otherwise, the world unchanged). Stepping after a load gives the same
 * results, bit for bit, as stepping the world that saved it. */
int pgx_world_state_size(pgx_world* w, size_t* bytes);
int pgx_world_save_state(pgx_world* w, void* buf, size_t bytes);
int pgx_world_load_state(pgx_world* w, const void* buf, size_t bytes);
/* Loads a state saved by a world whose colliders differ from this one's: the contact cache
 * kept through Scene::addCollider and Scene::clearColliders (src/Physecs.cpp:725-752; the
 * reference keys contactCache and triggerCache by entity and collider index,
 * include/Physecs/Physecs.h:112-125,165-171, and neither edit changes those of the other
 * colliders). Every cache entry, trigger pair and broadphase box goes to this world's
 * collider with the same world, entity and index within its entity; entries without one
 * are dropped. Poses and velocities stay as this world was created (from the descriptor).
 * The gear joints' state is loaded when both worlds have the same number of joints. The
 * last step's trigger pairs that lose a collider (all of them when this world has no
 * trigger colliders and no trigger-making contact filter) are written to `gone`, four
 * values each: entity0, index0, entity1, index1 (the lower entity first). The reference
 * reports them as exits at the next step (src/Physecs.cpp:545-551), this world does not:
 * the caller does. *n_gone gets their count; PGX_ERR_CAPACITY when it exceeds
 * gone_capacity, PGX_ERR_INVALID for a state of another number of worlds or a damaged one;
 * the world is unchanged on error. Load after pgx_world_set_contact_filter, which decides
 * whether the world keeps trigger lists. */
int pgx_world_load_state_matched(pgx_world* w, const void* buf, size_t bytes, int gone_capacity, uint32_t* gone,
                                 int32_t* n_gone);
/* In PGX_EXEC_COLORS, enabled != 0 launches each step's substep loop as one HIP graph:
 * captured every step (its launches follow the step's colours), the instantiated graph
 * updated in place when the launch sequence keeps its shape. Results are the same either
 * way. Steps with per-launch timing (pgx_world_enable_timing) launch directly. No
 * counterpart in the reference, whose substep loop is host code (src/Physecs.cpp:361-534). */
int pgx_world_set_graph(pgx_world* w, int enabled);

/* Launch shapes of the step. Results are the same for every value; only speed changes.
 * No counterpart in the reference (its CPU step has no launch shapes). */
typedef struct pgx_tuning {
    int solve_block;      /* threads per workgroup of the per-colour contact solve: 64, 128, 256 */
    int color_batch;      /* a run of two or more consecutive colours of at most this many
                             manifolds (and at most 1264 in all) is solved by one
                             single-workgroup launch that steps through them in order, its
                             bodies' velocities in LDS (0: one launch per colour; default 512) */
    int pairs_per_lane;   /* narrowphase pairs per lane: 1, 2, 4 */
    int color_rounds;     /* least graph-colouring rounds queued before the host reads back
                             whether colouring finished: 1..1024 */
    int eager_rows_max;   /* colours of at most this many manifolds issue all four point
                             slots' row loads at once instead of after the point count:
                             shorter latency, more bytes (>= 0) */
    int pair_search;      /* how the broadphase finds a step's pairs: 0 = a uniform grid of
                             the dynamic colliders (cells as large as the largest one; a
                             step whose fullest cell holds more than 512 colliders uses the
                             tree walk instead), 1 = a walk of the LBVH over escape links */
} pgx_tuning;
int pgx_world_set_tuning(pgx_world* w, const pgx_tuning* t);
int pgx_world_get_tuning(pgx_world* w, pgx_tuning* t);
/* Autotuner: times trial_steps steps (after one untimed step; the better of two runs)
 * from the world's current state for each value of one parameter at a time (color_batch,
 * solve_block, eager_rows_max, pairs_per_lane, color_rounds, in that order), keeps a value
 * that is at least 3% faster than the current choice, and restores the world's state
 * afterwards (poses, velocities, bounds, contact cache, gear angles, statistics), so the
 * world steps on exactly as it would have. The chosen values stay set
 * and are written to *chosen (may be NULL). A world in PGX_EXEC_COLORS or PGX_EXEC_WORLDS
 * whose worlds fit the per-world passes also has the faster of those two modes chosen
 * first (pgx_world_set_exec). */
int pgx_world_autotune(pgx_world* w, float dt, int trial_steps, pgx_tuning* chosen);

/* ---------------------------------------------------------------------------------
 * Serving: a request loop that keeps many independent scenes ("requests") resident in one
 * batched device world. pgx_server_submit queues a scene with a horizon (number of steps);
 * pgx_server_step steps every admitted request once, in the same launches. A request's
 * state is copied out right after its last step. At most every `admit_interval` steps
 * (and whenever nothing is running) the server retires finished requests and admits
 * queued ones, first come first served, while the HBM footprint of the admitted set stays
 * within the budget; until then a finished request keeps riding in the batch, unobserved.
 * Admission keeps the state of continuing requests (bodies, contact cache) on the device,
 * so a request's result is the same whether it ran alone or in the mix. Scenes with convex
 * or triangle meshes are not accepted by the server. */
typedef struct pgx_server pgx_server;

typedef struct pgx_server_config {
    double hbm_budget_bytes;   /* footprint budget of the admitted set (0: 64 GiB) */
    int max_worlds;            /* admitted requests at most (0: 8192) */
    int admit_interval;        /* steps between admission boundaries (0: 1) */
} pgx_server_config;

typedef struct pgx_server_step_info {
    int active;                /* requests advanced by this call (before their horizon) */
    int admitted, retired;     /* at this call's admission boundary */
    int finished;              /* requests whose last step this call was */
    int queued;                /* still waiting */
    long long bodies;          /* dynamic bodies of the advanced requests */
    double footprint_bytes;    /* estimated HBM footprint of the admitted set */
} pgx_server_step_info;

int pgx_server_create(const pgx_config* cfg, const pgx_server_config* scfg, pgx_server** out);
int pgx_server_destroy(pgx_server* s);
/* Queues a copy of the scene (and its joints); returns its request id. */
int pgx_server_submit(pgx_server* s, const pgx_scene_desc* scene, const pgx_joint_desc* joints, int n_joints,
                      int horizon_steps, int* request_id);
int pgx_server_step(pgx_server* s, float dt, pgx_server_step_info* info);
/* 1 when the request finished (its state is available), 0 while queued or running. */
int pgx_server_done(pgx_server* s, int request_id);
/* the batched world's substep loop as one HIP graph launch per step (pgx_world_set_graph) */
int pgx_server_set_graph(pgx_server* s, int enabled);
/* launch shapes of the batched world (pgx_world_set_tuning), e.g. from pgx_world_autotune
 * on a world of the same request mix */
int pgx_server_set_tuning(pgx_server* s, const pgx_tuning* t);
/* Final state of a finished request in its descriptor's order (pos/quat per actor,
 * vel/angvel per dynamic body); releases it. */
int pgx_server_take_result(pgx_server* s, int request_id, float* pos, float* quat, float* vel, float* angvel);
/* Estimated HBM footprint of one request (bytes). */
int pgx_server_footprint(const pgx_scene_desc* scene, double* bytes);

#ifdef __cplusplus
}
#endif

#endif /* PHYSECS_GPU_H */
