/* natsel_gpu.h - C ABI of the MI355X (gfx950) rigid-body stepper.
 *
 * The reference engine has no FFI: its seam is the public World API. Each entry point
 * below stands in for one World method so the GPU stepper is a drop-in for the CPU
 * stepper (citations are to the reference sources):
 *
 *   nsl_world_create / destroy     World::new / drop of the World  src/simulation/world_runtime.rs:1162
 *   nsl_world_set_options          public World fields          src/simulation/world_runtime.rs:1243-1304
 *   nsl_create_body                World::create_body           src/simulation/world_runtime.rs:1506
 *   nsl_default_body_def           BodyDef::default             src/types/body_def.rs:80-104
 *   nsl_default_material           ShapeDef::default            src/types/shape_def.rs:114-135
 *   nsl_create_box                 create_hull_shape(make_box_hull)  src/simulation/shape_creation.rs:559,
 *                                                               src/collision/hull.rs:854
 *   nsl_create_sphere              World::create_sphere_shape   src/simulation/shape_creation.rs:504
 *   nsl_create_capsule             World::create_capsule_shape  src/simulation/shape_creation.rs:522
 *   nsl_world_step                 World::step(dt, substeps)    src/simulation/solver_integration.rs:807
 *   nsl_world_get_bodies           body_transform / body_linear_velocity / body_angular_velocity
 *                                                               src/simulation/body_api.rs:186-262
 *   nsl_world_state_hash           World::world_state_hash      src/simulation/body_queries.rs:432
 *   nsl_world_get_contact_colors   constraint-graph colour of each contact (graph_tests.rs:152-172)
 *   nsl_world_get_contact_shapes   the shape pair of each contact id (world_runtime.rs:517-555)
 *   nsl_world_body_location        Body::set_index / local_index / island_id; body_is_awake
 *                                                               src/simulation/body_api.rs:836-839
 *   nsl_world_island_summary       island pool and solver-set counts (island_split_tests.rs:46-117)
 *   nsl_world_check_islands        the reference's index validation helpers
 *   nsl_time_of_impact             time_of_impact               src/collision/time_of_impact.rs:406-598
 *   nsl_collide                    collide_hulls / collide_spheres / collide_capsule_and_sphere /
 *                                  collide_capsules / collide_hull_and_sphere / collide_hull_and_capsule
 *                                                               src/collision/convex_manifold.rs:1419-1437,
 *                                                               src/collision/convex_manifold_shapes.rs:21-797
 *   nsl_world_get_profile          World::profile               src/types/world_def.rs:158-184
 *   nsl_world_autotune             the step kernel's launch shape by timing (no counterpart: the reference
 *                                  sizes its stage blocks by fixed rules, solver_stages.rs:490-525)
 *   nsl_world_*kernel_*            per-kernel device timing (no counterpart: the reference times phases
 *                                  on the host, solver_stages.rs:551-586)
 *   nsl_batch_*                    many Worlds at once: create_body / create_*_shape / step /
 *                                  world_state_hash per world, as above (no batch type in the reference)
 *   nsl_batch_*tuning, _autotune   the batched step's launch shape (no counterpart, as nsl_world_autotune)
 *
 * Conventions: every call returns 0 on success or a negative status; on failure
 * nsl_last_error() returns a message. Pointers are plain host memory owned by the
 * caller and are copied during the call. One world per host thread; calls on a world
 * are ordered on the world's own HIP stream and return after their device work (and
 * any copy back) has finished. Rotations are quaternions (x, y, z, s).
 */
#ifndef NATSEL_GPU_H
#define NATSEL_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NSL_OK 0
#define NSL_ERR_INVALID (-1)   /* bad argument or handle */
#define NSL_ERR_DEVICE (-2)    /* HIP runtime error */
#define NSL_ERR_INTERNAL (-3)  /* other failure */

#define NSL_BODY_STATIC 0
#define NSL_BODY_KINEMATIC 1
#define NSL_BODY_DYNAMIC 2

typedef struct nsl_world nsl_world;

typedef struct nsl_options {
  float gravity[3];                  /* default (0, -10, 0) */
  float contact_hertz;               /* 30 */
  float contact_damping_ratio;       /* 10 */
  float contact_speed;               /* 3 m/s */
  float max_linear_speed;            /* 400 m/s */
  float restitution_threshold;       /* 1 m/s */
  int32_t restitution_iterations;    /* 2 */
  int32_t enable_sleep;              /* 1 */
  int32_t enable_continuous;         /* 1: fast bodies are swept (continuous.rs) */
  int32_t enable_warm_starting;      /* 1 */
  int32_t enable_speculative;        /* 1 */
  int32_t solver_blocks;             /* workgroups of the step kernel; 0 = sized from the largest phase */
  int32_t enable_restitution_propagation; /* 0 (world_runtime.rs:1055) */
  int32_t small_colour_items;        /* 128: consecutive colours of at most this many items run on one
                                        workgroup with workgroup barriers, one grid barrier after the
                                        run (0: every colour behind a grid barrier); results do not change */
  int32_t stage_records;             /* 0: 1 = the step kernel copies each lane's constraint record of the
                                        next colour to LDS while the current colour runs; results do not change */
  int32_t plain_step_launch;         /* 0: the step kernel starts through hipLaunchCooperativeKernel (the runtime
                                        guarantees that every workgroup is resident); 1 = an ordinary launch of
                                        the same grid, sized to the occupancy limit, for tools that cannot
                                        follow cooperative launches; results do not change */
} nsl_options;

typedef struct nsl_body_def {
  int32_t type;                      /* NSL_BODY_* */
  double position[3];
  float rotation[4];
  float linear_velocity[3];
  float angular_velocity[3];
  float linear_damping, angular_damping, gravity_scale;
  int32_t enable_sleep;
  int32_t is_bullet;                 /* swept against every body type, after the other fast bodies */
  float safety_factor;               /* 0.5: fast (swept) when a step moves it more than this times
                                        its smallest extent (body_def.rs:68, continuous.rs:121-151) */
  float sleep_threshold;             /* 0.05 m/s: sleep velocity threshold (body_def.rs:67) */
  int32_t motion_locks;              /* NSL_LOCK_* bits (MotionLocks, body_def.rs:24-33) */
  int32_t is_awake;                  /* 1. A dynamic or kinematic body with is_awake 0 and enable_sleep 1
                                        would start in a sleeping set of its own: refused (NSL_ERR_INVALID) */
  int32_t is_enabled;                /* 1. 0 (the disabled set) is refused (NSL_ERR_INVALID) */
  int32_t allow_fast_rotation;       /* 0: the angular speed clamp applies (solver_integration.rs:377-426) */
  int32_t enable_contact_recycling;  /* 1: contacts of this body may skip the narrow phase when their
                                        relative pose barely moved (collide_pass.rs:606-687) */
} nsl_body_def;

/* MotionLocks as body flag bits (world_runtime.rs:87-92). */
#define NSL_LOCK_LINEAR_X 0x1
#define NSL_LOCK_LINEAR_Y 0x2
#define NSL_LOCK_LINEAR_Z 0x4
#define NSL_LOCK_ANGULAR_X 0x8
#define NSL_LOCK_ANGULAR_Y 0x10
#define NSL_LOCK_ANGULAR_Z 0x20

/* A shape's ShapeDef (src/types/shape_def.rs:90-135) with its base SurfaceMaterial. Fill it with
   nsl_default_material before setting fields: a zero mask filters every pair out. */
typedef struct nsl_material {
  float density, friction, restitution, rolling_resistance; /* 1000, 0.6, 0, 0 */
  uint64_t filter_category_bits;      /* 1 (Filter, shape_def.rs:18-24; shape_runtime.rs:285-292) */
  uint64_t filter_mask_bits;          /* every bit */
  int32_t filter_group_index;         /* 0 */
  int32_t is_sensor;                  /* 0. A sensor creates no contacts and takes no part in
                                         continuous collision (contact_creation.rs:187,
                                         continuous.rs:398-401); sensor events are not produced */
  int32_t enable_speculative_contact; /* 1 */
  int32_t enable_hit_events;          /* 0 (flag kept on the contact; events are not produced) */
} nsl_material;

typedef struct nsl_profile {
  float pairs_ms, collide_ms, state_pass_ms, solve_ms, finalize_ms, step_ms;
  int32_t contacts, touching, colors_used, overflow;
  int32_t fast_bodies, sat_calls, sat_hits, recycled;
} nsl_profile;

/* Launch classes of nsl_kernel_stats.ms / launches (HIP events around each launch):
   NSL_K_SOLVE is the step kernel, NSL_K_PREPARE the constraint preparation before it and
   NSL_K_FINALIZE the impulse store and body finalize after it. */
enum { NSL_K_INIT = 0, NSL_K_COLLIDE, NSL_K_SOLVE, NSL_K_CONTINUOUS, NSL_K_PREPARE, NSL_K_FINALIZE, NSL_K_CLASS_COUNT };

/* Phases of the step kernel (NSL_K_SOLVE), timed with the device wall clock at its grid
 * barriers; each phase's time includes the barrier that ends it. NSL_P_PREPARE and
 * NSL_P_STORE_FINALIZE stay 0 (those are the NSL_K_PREPARE / NSL_K_FINALIZE launches). */
enum {
  NSL_P_PREPARE = 0, NSL_P_INTEGRATE_VELOCITIES, NSL_P_WARM_START, NSL_P_PUSH, NSL_P_INTEGRATE_POSITIONS,
  NSL_P_RELAX, NSL_P_RESTITUTION, NSL_P_OVERFLOW, NSL_P_STORE_FINALIZE, NSL_P_CLASS_COUNT
};

/* Device time (only while timing is on) and work counters since the last reset. */
typedef struct nsl_kernel_stats {
  double ms[NSL_K_CLASS_COUNT];
  int64_t launches[NSL_K_CLASS_COUNT];
  double phase_ms[NSL_P_CLASS_COUNT];
  int64_t phase_count[NSL_P_CLASS_COUNT];
  int64_t steps, bodies, collided, constraints, points, overflow_points, recycled, sat_calls, sat_hits, fast_bodies;
  int64_t barriers;      /* grid barriers executed by the step kernel */
  int64_t solver_blocks; /* workgroups of the last step kernel launch */
  int64_t pair_tests;    /* candidate shape pairs tested by the device pair finder */
  double host_ms[4];     /* host wall time: pair update, collide (incl. wait), state pass, solve+finalize (incl. wait) */
  int64_t joints;        /* joints solved, summed over steps */
  int64_t toi_calls;     /* continuous collision: time-of-impact queries */
  int64_t toi_hits;      /* continuous collision: fast bodies moved back to an impact */
} nsl_kernel_stats;

/* One time-of-impact query (TOIInput, src/types/queries.rs:201-207): two proxies of 1 to 8
   support points with a radius, and their sweeps: local centre, c1, c2, q1 (x, y, z, s), q2. */
typedef struct nsl_toi_input {
  float points_a[8][3], points_b[8][3];
  int32_t count_a, count_b;
  float radius_a, radius_b;
  float sweep_a[17], sweep_b[17];
  float max_fraction;
} nsl_toi_input;

enum { NSL_TOI_UNKNOWN = 0, NSL_TOI_FAILED, NSL_TOI_OVERLAPPED, NSL_TOI_HIT, NSL_TOI_SEPARATED };

/* TOIOutput (src/types/queries.rs:224-234). */
typedef struct nsl_toi_output {
  int32_t state;
  float point[3], normal[3];
  float fraction, distance;
  int32_t distance_iterations, push_back_iterations, root_iterations;
} nsl_toi_output;

/* A revolute joint (RevoluteJointDef, src/types/joint_defs.rs:30-58, :154-167). Frames
   are (px, py, pz, qx, qy, qz, qs) in each body's frame; the hinge axis is frame A's z. */
typedef struct nsl_revolute_joint_def {
  int32_t body_a, body_b;
  float local_frame_a[7];
  float local_frame_b[7];
  float constraint_hertz;            /* 60 */
  float constraint_damping_ratio;    /* 2 */
  int32_t collide_connected;         /* 0: the bodies' contacts are destroyed and no new one forms */
  int32_t enable_spring;             /* 0 */
  float hertz, damping_ratio, target_angle;
  int32_t enable_limit;              /* 0 */
  float lower_angle, upper_angle;
  int32_t enable_motor;              /* 0 */
  float max_motor_torque, motor_speed;
} nsl_revolute_joint_def;

const char* nsl_last_error(void);
void nsl_default_options(nsl_options* out);
void nsl_default_revolute_joint_def(nsl_revolute_joint_def* out);
/* BodyDef::default (src/types/body_def.rs:80-104): a static body at the origin. */
void nsl_default_body_def(nsl_body_def* out);
/* ShapeDef::default (src/types/shape_def.rs:114-135) with SurfaceMaterial::default_material. */
void nsl_default_material(nsl_material* out);

/* time_of_impact (src/collision/time_of_impact.rs:406-598) for n queries on the device, one
   lane each; host arrays in and out. */
int nsl_time_of_impact(int device, const nsl_toi_input* in, nsl_toi_output* out, int32_t n);

/* ---------------------------------------------------------------------------------
 * The narrow phase on its own. Each of the reference's collide routines takes two shapes,
 * the transform of B in A's frame, a caller buffer for the manifold and the contact's cache
 * (collide_hulls, convex_manifold.rs:1419-1437; collide_spheres, collide_capsule_and_sphere,
 * collide_capsules, collide_hull_and_sphere, collide_hull_and_capsule,
 * convex_manifold_shapes.rs:21-797). nsl_collide runs n such calls on the device, one lane
 * each, in the pair order of the register table (contact_creation.rs:88-102): hull-hull,
 * hull-sphere, hull-capsule, capsule-sphere, capsule-capsule, sphere-sphere. The results
 * equal the step's own narrow phase bit for bit. */
#define NSL_SHAPE_CAPSULE 0 /* ShapeType values (src/types/shape_def.rs:68-82) */
#define NSL_SHAPE_HULL 3
#define NSL_SHAPE_SPHERE 5

typedef struct nsl_shape_geometry {
  int32_t type;    /* NSL_SHAPE_* */
  float params[7]; /* hull (a box): half extents hx, hy, hz; sphere: centre xyz, radius;
                      capsule: centre1 xyz, centre2 xyz, radius */
  float frame[7];  /* hull only: where the box sits in its shape frame, p xyz and q xyzs
                      (make_transformed_box_hull, src/collision/hull.rs:754-852) */
} nsl_shape_geometry;

/* The cache a contact carries from call to call, zeroed when empty: SATCache for hull-hull
   (src/types/collision_types.rs:97-121: separation, SeparatingFeature type, index_a,
   index_b, hit), SimplexCache for hull-sphere and hull-capsule (src/types/queries.rs: metric,
   count, index_a[0..count), index_b[0..count)); unused by the other pairs. */
typedef struct nsl_collide_cache {
  float value;                    /* SATCache.separation or SimplexCache.metric */
  int32_t kind;                   /* SATCache.type or SimplexCache.count (0..3) */
  uint8_t index_a[4], index_b[4]; /* SATCache: [0] only */
  int32_t hit;                    /* SATCache.hit (set by the call) */
} nsl_collide_cache;

typedef struct nsl_collide_query {
  nsl_shape_geometry a, b;
  float transform_b_to_a[7]; /* p xyz, q xyzs */
  nsl_collide_cache cache;
} nsl_collide_query;

/* LocalManifold with capacity 4 (MAX_MANIFOLD_POINTS): normal and points in A's frame,
   separations, feature ids (src/collision/manifold.rs:58-94), and the cache after the call. */
typedef struct nsl_local_manifold {
  float normal[3];
  int32_t point_count;
  float point[4][3];
  float separation[4];
  uint32_t feature_id[4];
  nsl_collide_cache cache;
} nsl_local_manifold;

/* n collide calls on the device; host arrays in and out. NSL_ERR_INVALID for a pair outside
   the register table's order or a cache whose indices do not fit its shapes (SAT face index
   above 5, SAT edge index odd or above 22, simplex count above 3 or a vertex index beyond the
   shape's support points). */
int nsl_collide(int device, const nsl_collide_query* queries, nsl_local_manifold* out, int32_t n);

/* nsl_collide without any allocation. No counterpart in the reference: for callers that
   cannot allocate device memory (code under HIP graph capture, tight serving loops). The
   queries and manifolds are staged in a fixed device array of the library (about 2.2 MiB on
   each device it runs on, held until the process ends) instead of buffers allocated per
   call; results equal nsl_collide's. At most 8192 queries per call (NSL_ERR_INVALID above).
   Every call on one device shares that device's array: calls that overlap on one device
   (from several host threads) race. */
int nsl_collide_noalloc(int device, const nsl_collide_query* queries, nsl_local_manifold* out, int32_t n);

int nsl_world_create(int device, const nsl_options* options, nsl_world** out);
int nsl_world_destroy(nsl_world* world);
int nsl_world_set_options(nsl_world* world, const nsl_options* options);

int nsl_create_body(nsl_world* world, const nsl_body_def* def, int32_t* out_body);
int nsl_create_box(nsl_world* world, int32_t body, const float half_extents[3], const nsl_material* m, int32_t* out_shape);
int nsl_create_sphere(nsl_world* world, int32_t body, const float center[3], float radius, const nsl_material* m,
                      int32_t* out_shape);
int nsl_create_capsule(nsl_world* world, int32_t body, const float c1[3], const float c2[3], float radius,
                       const nsl_material* m, int32_t* out_shape);

/* World::create_revolute_joint (src/simulation/revolute_joint.rs:149-174; joint.rs:164-375). */
int nsl_create_revolute_joint(nsl_world* world, const nsl_revolute_joint_def* def, int32_t* out_joint);
/* Accumulated impulses of a joint: linear xyz, perpendicular xy, spring, motor, lower,
   upper (RevoluteJointSim, revolute_joint.rs:27-57) and its graph colour (-1: not in the
   graph). */
int nsl_world_get_joint(nsl_world* world, int32_t joint, float impulses[9], int32_t* color);
int nsl_world_joint_count(nsl_world* world, int32_t* out);
int nsl_world_contact_count(nsl_world* world, int32_t* out);

/* World::save_state / World::restore_state (src/simulation/world_state.rs:340, :439): a
 * snapshot holds everything that decides later steps (bodies, contacts and their caches,
 * the pair set, joints, ids, colours). Restoring needs the bodies, shapes and joints the
 * world held when the snapshot was taken. */
typedef struct nsl_snapshot nsl_snapshot;
int nsl_world_save_state(nsl_world* world, nsl_snapshot** out);
int nsl_world_restore_state(nsl_world* world, const nsl_snapshot* snapshot);
int nsl_snapshot_destroy(nsl_snapshot* snapshot);

int nsl_world_step(nsl_world* world, float dt, int32_t sub_steps);
int nsl_world_step_n(nsl_world* world, float dt, int32_t sub_steps, int32_t steps);

int nsl_world_body_capacity(nsl_world* world, int32_t* out);
/* Arrays sized by nsl_world_body_capacity: pos[3n] (f64), rot[4n], v[3n], w[3n]. */
int nsl_world_get_bodies(nsl_world* world, double* pos, float* rot, float* v, float* w);
int nsl_world_state_hash(nsl_world* world, uint64_t* out);
int nsl_world_contact_capacity(nsl_world* world, int32_t* out);
/* Arrays sized by nsl_world_contact_capacity. */
int nsl_world_get_contact_colors(nsl_world* world, int32_t* color, int32_t* local, int32_t* alive);
/* Per contact id: its two shape ids (Contact::shape_id_a/_b, world_runtime.rs:517-555), -1 when
   the id is free. */
int nsl_world_get_contact_shapes(nsl_world* world, int32_t* shape_a, int32_t* shape_b);
/* A body's solver set (0 static, 2 awake, 3 and up sleeping), its index in that set and its
   island id (-1 for a static body): Body::set_index / local_index / island_id
   (world_runtime.rs:301-343), read by the reference's sleep tests (sleep_tests.rs:33-99);
   body_is_awake (body_api.rs:836-839) is set == 2. */
int nsl_world_body_location(nsl_world* world, int32_t body, int32_t* set, int32_t* local, int32_t* island);
/* out[5]: island slots, live islands, awake islands, sleeping sets, the island to split at the
   next step (-1: none) (island_split_tests.rs:46-117 reads the first two). */
int nsl_world_island_summary(nsl_world* world, int64_t* out);
/* *out = 0 when every island, solver-set and contact-list record agrees with the bodies' and
   contacts' own records, else the number of the first failed check (the reference's
   validate_graph_indices / validate_set_contacts / validate_islands, constraint_graph.rs:370-560). */
int nsl_world_check_islands(nsl_world* world, int32_t* out);
int nsl_world_get_profile(nsl_world* world, nsl_profile* out);
/* on: bit 0 times every launch with an event pair (nsl_kernel_stats.ms / launches); bit 1
   (with bit 0) also runs the step kernel's phase clock, a wall-clock read per grid barrier
   and a download of the marks after each step (phase_ms / phase_count). Results do not
   change. */
int nsl_world_set_kernel_timing(nsl_world* world, int32_t on);

/* The launch shape the autotuner picked (nsl_world_autotune). */
typedef struct nsl_tuning {
  int32_t solver_blocks, small_colour_items, stage_records; /* now set in the world's options */
  double step_kernel_us;    /* average step-kernel time of that shape over the trial steps */
  double previous_us;       /* the same for the options the world had before the call */
  int32_t candidates;       /* shapes tried */
} nsl_tuning;
/* Picks the step kernel's launch shape for this world as it is now: from a snapshot, runs
   `steps` steps of (dt, sub_steps) under each candidate (workgroups around the size the
   widest colour needs, the small-colour threshold 0 / 128 / 256, record staging off / on),
   times the step kernel with events, checks that every candidate ends on the same state
   hash (the shape never changes results; NSL_ERR_INTERNAL otherwise), restores the snapshot
   and keeps the fastest shape in the world's options. No reference counterpart: the
   reference sizes its stage blocks by fixed rules (solver_stages.rs:490-525). */
int nsl_world_autotune(nsl_world* world, float dt, int32_t sub_steps, int32_t steps, nsl_tuning* out);
/* Per-colour residuals of the coloured contact solver. No reference counterpart: a
   convergence diagnostic of the graph-coloured relax stage (contact_solver.rs:927-1097).
   While on, each step ends with one more relax sweep over the coloured contact constraints,
   colour by colour in the solver's order, from the state the step's solver ends in (after
   the last relax and restitution), on a copy of the body states: the world's state and the
   step's results are unchanged. Each wavefront reduces its constraints' impulse changes with
   lane shuffles; per colour: the sum over points of |normal impulse change| (wavefront
   partials added in slot order in double), the largest such change, the largest length of
   a friction impulse change, and the number of constraints. Joints and the overflow colour
   are not part of the sweep. The results are the last step's; zeros before the first step
   with residuals on. */
#define NSL_GRAPH_COLOURS 23 /* colours 0..22; 23 is the overflow colour (constraint_graph.rs:20) */
typedef struct nsl_colour_residual {
  double sum_normal;
  float max_normal, max_friction;
  int32_t constraints, reserved;
} nsl_colour_residual;
int nsl_world_set_residuals(nsl_world* world, int32_t on);
int nsl_world_get_residuals(nsl_world* world, nsl_colour_residual* out, int32_t n);
int nsl_world_reset_kernel_stats(nsl_world* world);
int nsl_world_get_kernel_stats(nsl_world* world, nsl_kernel_stats* out);

/* ---------------------------------------------------------------------------------
 * Batches of independent worlds. Each world is built and stepped exactly as a World of
 * its own (the calls mirror the World methods above, applied to world `world`), and
 * its state hash equals the hash the same world reaches alone. All worlds of a batch
 * share the options and step together; worlds, bodies and shapes can be added until the
 * first nsl_batch_step. A world holds at most 256 dynamic bodies; a world whose contact
 * capacity runs out stops stepping and is reported by nsl_batch_get_stats. Sleep and
 * continuous collision are not run. */
typedef struct nsl_batch nsl_batch;

typedef struct nsl_batch_stats {
  double last_step_ms;      /* device time of the last nsl_batch_step call */
  int64_t body_steps;       /* dynamic bodies x steps since creation */
  int64_t bodies;           /* dynamic bodies in the batch */
  int64_t contact_capacity; /* contact slots over all worlds */
  int32_t error_worlds;     /* worlds stopped by a full capacity */
  int32_t first_error;      /* 1: contacts full, 2: overflow colour full, 3: new pairs of a step over 2048 */
  int64_t contacts_collided;    /* since creation, over all worlds and steps */
  int64_t constraints;          /* coloured contact constraints prepared */
  int64_t points;               /* their manifold points */
  int64_t overflow_constraints; /* overflow-colour constraints prepared */
} nsl_batch_stats;

int nsl_batch_create(int device, const nsl_options* options, nsl_batch** out);
int nsl_batch_destroy(nsl_batch* batch);
int nsl_batch_add_world(nsl_batch* batch, int32_t* out_world);
int nsl_batch_create_body(nsl_batch* batch, int32_t world, const nsl_body_def* def, int32_t* out_body);
int nsl_batch_create_box(nsl_batch* batch, int32_t world, int32_t body, const float half_extents[3], const nsl_material* m,
                         int32_t* out_shape);
int nsl_batch_create_sphere(nsl_batch* batch, int32_t world, int32_t body, const float center[3], float radius,
                            const nsl_material* m, int32_t* out_shape);
int nsl_batch_create_capsule(nsl_batch* batch, int32_t world, int32_t body, const float c1[3], const float c2[3], float radius,
                             const nsl_material* m, int32_t* out_shape);
/* `steps` steps of every world, with no host work between them. */
int nsl_batch_step(nsl_batch* batch, float dt, int32_t sub_steps, int32_t steps);
int nsl_batch_world_count(nsl_batch* batch, int32_t* out);
int nsl_batch_body_count(nsl_batch* batch, int32_t world, int32_t* out);
/* out[world_count]: each world's World::world_state_hash. */
int nsl_batch_state_hashes(nsl_batch* batch, uint64_t* out);
/* Arrays sized by nsl_batch_body_count(world), as nsl_world_get_bodies. */
int nsl_batch_get_bodies(nsl_batch* batch, int32_t world, double* pos, float* rot, float* v, float* w);
/* Contact slots of a world (0 before the first step) and, per slot, its constraint-graph
 * colour (-1 when not in the graph), local index and liveness, as nsl_world_get_contact_colors. */
int nsl_batch_contact_capacity(nsl_batch* batch, int32_t world, int32_t* out);
int nsl_batch_get_contact_colors(nsl_batch* batch, int32_t world, int32_t* color, int32_t* local, int32_t* alive);
/* Per contact slot of a world: its shape ids (primary first; -1 for a free slot). */
int nsl_batch_get_contact_shapes(nsl_batch* batch, int32_t world, int32_t* shape_a, int32_t* shape_b);
/* Per shape of a world (shape count = shapes created in it): fat AABB as lower x, y, z,
 * upper x, y, z (shape_runtime.rs:398-408), and whether its proxy moved since the last
 * pair update. */
int nsl_batch_get_shape_bounds(nsl_batch* batch, int32_t world, float* fat, uint8_t* moved);
int nsl_batch_get_stats(nsl_batch* batch, nsl_batch_stats* out);

/* Launch shape of nsl_batch_step; results do not depend on it. */
typedef struct nsl_batch_tuning {
  int32_t collide_blocks;   /* workgroups of the collide launch (it strides over the live contacts);
                               0 = as many as the device keeps resident at once */
  int32_t use_graph;        /* 1: a call replays a captured HIP graph of 8 steps as often as it fits,
                               the remaining steps launched directly */
  int32_t candidates;       /* nsl_batch_autotune: shapes timed */
  int32_t reserved;
  double step_ms;           /* nsl_batch_autotune: the chosen shape's device time per step */
  double previous_ms;       /* nsl_batch_autotune: the same for the shape in use before the call */
} nsl_batch_tuning;
int nsl_batch_set_tuning(nsl_batch* batch, const nsl_batch_tuning* tuning);
int nsl_batch_get_tuning(nsl_batch* batch, nsl_batch_tuning* out);
/* Picks the launch shape by timing: from a device snapshot of the whole batch, runs `steps`
   steps of (dt, sub_steps) under each candidate (collide workgroups 1, 2, 4, 8 or 16 times the
   resident count, graph replay off / on), checks that every candidate ends on the same world hashes
   (NSL_ERR_INTERNAL otherwise), restores the snapshot and keeps the fastest shape. The
   snapshot holds a second copy of the batch's device arrays during the call. No reference
   counterpart (the reference has no batch type). */
int nsl_batch_autotune(nsl_batch* batch, float dt, int32_t sub_steps, int32_t steps, nsl_batch_tuning* out);

/* Differentiable rollout of a batch. The reference has no differentiable path; these
 * entry points are for gradient-based callers (trajectory optimisation, system
 * identification) of the batched stepper.
 *
 * nsl_batch_record steps like nsl_batch_step while recording a tape of every step (at most
 * max_constraints <= 64 contact constraints and 64 dynamic bodies per world).
 * nsl_batch_adjoint then returns, per world, the loss |com - target|^2 of the dynamic bodies'
 * centre of mass after the recorded steps (targets: [worlds][3]) and, per dynamic body
 * (arrays in the batch's body order, world after world), the gradient of its world's loss
 * with respect to the body's linear and angular velocity at the first recorded step
 * ([bodies][3] each) and its shape's friction ([bodies]). The differentiated model holds the
 * recorded active set, colour order and contact geometry fixed, lets each point's separation
 * follow the body centres and carries warm-start impulses between steps as the stepper does.
 * status per world: 0 ok; 1 a recorded step lies outside the model (overflow colour,
 * restitution, rolling resistance, capacity); 2 the adjoint's recompute departed from the
 * recording.
 * nsl_batch_model_losses runs the same model in double precision from the recorded first
 * state for n cases, each with one parameter of one world perturbed by deltas[i]
 * (params[i] = awake body index in its world * 7 + {0..2: v, 3..5: w, 6: friction}), and
 * returns the loss of each case: central finite differences of the adjoint's model. With
 * frozen != 0 every clamp, cone and separation branch is taken from the branch record of the
 * last nsl_batch_adjoint (the model with the recorded active set, which the adjoint
 * differentiates); with frozen = 0 the model decides its own branches. */
int nsl_batch_record(nsl_batch* batch, float dt, int32_t sub_steps, int32_t steps, int32_t max_constraints);
int nsl_batch_adjoint(nsl_batch* batch, const float* targets, float* loss, float* grad_v, float* grad_w, float* grad_friction,
                      int32_t* status);
int nsl_batch_model_losses(nsl_batch* batch, const int32_t* worlds, const int32_t* params, const double* deltas, int32_t n,
                           const float* targets, int32_t frozen, double* out);
/* Device time of the last nsl_batch_record and nsl_batch_adjoint calls. */
int nsl_batch_rollout_times(nsl_batch* batch, double* record_ms, double* adjoint_ms);

#ifdef __cplusplus
}
#endif

#endif /* NATSEL_GPU_H */
