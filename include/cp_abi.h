/* cp_abi.h — C ABI of the MI355X (gfx950) contact-pipeline library
 * (constraintphysics_amd/libcp_hip.so).
 *
 * The boundary sits where NathanMacLeod/ConstraintPhysics' PhysicsEngine
 * (ConstraintPhysics/src/PhysicsEngine.h:60-165) hands work to its per-step
 * pipeline.  No C++ or torch types cross it: plain pointers, sizes and int
 * status codes.  Each entry point names the reference interface it replaces.
 *
 * Conventions
 *   - Every function returns cp_status (0 = OK).  No exceptions cross the ABI.
 *   - Bodies are indexed densely from 0 in creation order; the reference body
 *     id is index + 1 (PhysicsEngine.cpp:487, next_id starts at 1).
 *   - Body state rows are 13 doubles: com xyz, orientation (r, i, j, k),
 *     velocity xyz, angular velocity xyz.
 *   - Contact cache rows: 6 int64 (id1, id2, memory_life, is_live,
 *     static_ready, position in edge), 2 uint64 MagicID (cID, bID), 21 doubles
 *     (contact point in b1 and b2 frames, normal in b1 frame, pen depth,
 *     restitution, friction coefficient, normal impulse, 2 friction impulses,
 *     friction basis u and w).  cID halves encode primitives as
 *     body_id * 16 + primitive index.
 *   - Bodies of one scene must be added contiguously; scenes never interact.
 */
#ifndef CP_ABI_H
#define CP_ABI_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
	CP_OK = 0,
	CP_ERR_INVALID = 1,   /* bad argument */
	CP_ERR_OOM = 2,       /* device allocation failed */
	CP_ERR_DEVICE = 3,    /* HIP runtime error */
	CP_ERR_CAPACITY = 4,  /* caller buffer too small (required size returned) */
	CP_ERR_STATE = 5,     /* call not valid in the current world state */
	CP_ERR_NONFINITE = 6, /* cp_world_step: a dynamic body's pose or velocity became NaN or infinite
	                         (the step completed; cp_world_scene_flags names the scenes).
	                         cp_world_get_scene_state: that scene holds such a body. */
	CP_ERR_IO = 7         /* a file could not be opened, read or written, or is not a valid file of
	                         its kind (wrong magic, sizes or checksum) */
} cp_status;

/* CP_PGS_COLORED: contacts only, graph-coloured per scene (FP64 or split).
 * CP_PGS_REFERENCE_ORDER: the reference's island search and sequential row
 * order (PhysicsEngine.cpp:1536-1613, ConstraintSolver.cpp:80-213), joints and
 * holonomic systems included; FP64.  CP_PGS_ISLAND_COLORED: the same islands,
 * rows and holonomic solve, with each island's edges graph-coloured so that a
 * wave sweeps an island in parallel; FP64. */
typedef enum { CP_PGS_COLORED = 0, CP_PGS_REFERENCE_ORDER = 1, CP_PGS_ISLAND_COLORED = 2 } cp_pgs_mode;

/* CP_PREC_FP64: rows, impulses and velocity deltas in FP64 (the reference's
 * arithmetic).  CP_PREC_SPLIT: contact rows derived in FP64 and stored in a
 * compact FP32 record, PGS sweeps, impulses within the solve and (in the
 * batched coloured mode) velocity deltas in FP32; body state, integration,
 * contact cache, narrowphase, joint rows and the holonomic block solve stay
 * FP64. */
typedef enum { CP_PREC_FP64 = 0, CP_PREC_SPLIT = 1 } cp_precision;

/* Engine parameters; defaults of PhysicsEngine.h:174-216 via cp_config_default. */
typedef struct {
	double dt;               /* setStep_time (PhysicsEngine.h:98) */
	double gravity[3];       /* setGravity */
	int vel_iters;           /* setPGSIterations(n_vel, n_pos, n_holonomic) */
	int pos_iters;
	int hol_iters;
	int substeps;            /* setSubstepCount */
	double cfm;              /* setGlobalConstraintForceMixing */
	double contact_hardness; /* contact_pos_correct_hardness (PhysicsEngine.h:216) */
	int contact_life;        /* contact_life (PhysicsEngine.h:215) */
	int gyro_ticks;          /* setAngleVelUpdateTickCount */
	int pgs_mode;            /* cp_pgs_mode */
	int solver_threads;      /* threads per scene workgroup in colored mode for scenes whose velocity
	                            deltas exceed LDS (64..1024, a multiple of 64) */
	int diagnostics;         /* bit 0: measure the FP64 PGS residual of every step's last substep;
	                            bit 1: record the coloured sweep schedule (cp_world_last_islands) */
	int defect_fixes;        /* bit 0: drop non-finite manifold points (the reference's clip
	                            intersection divides by n.d, which rounds to 0 for an edge parallel to
	                            the clipping edge, CollisionDetect.cpp:444-467); bit 1: a body whose
	                            bounding box holds a NaN or an infinity collides with nothing (the
	                            reference's AABB::intersects, AABB.h:10-14, lets a NaN box overlap every
	                            box, so one diverged body would pair with its whole scene).  Default 3;
	                            0 = bug-compatible */
	int precision;           /* cp_precision: solve precision of CP_PGS_COLORED and of the contact rows
	                            of CP_PGS_ISLAND_COLORED (joint rows, the holonomic block solve and the
	                            island solve's velocity deltas stay FP64 there); CP_PGS_REFERENCE_ORDER,
	                            the parity mode, is always FP64 */
	int pgs_class_bodies[3]; /* CP_PGS_COLORED size classes: scenes of at most [0], [1], [2] bodies
	                            (ascending, [2] <= 3072) keep their velocity deltas in LDS; larger
	                            scenes use global memory.  Defaults 192, 768, 3072. */
	int pgs_class_threads[3]; /* threads per scene workgroup of each LDS class (64..1024, a multiple of
	                            64).  Defaults 64, 128, 256.  Neither setting changes results. */
	int graph;               /* 1: in CP_PGS_COLORED mode without the residual diagnostics, the
	                            substep loop (rows, PGS, write-back, integration) is captured once in
	                            a HIP graph and replayed each step (recaptured when buffers, sizes or
	                            settings change); the per-solve PGS time is then not measured.
	                            Default 0.  Does not change results. */
	int fused;               /* 1: in CP_PGS_COLORED mode without the residual diagnostics and the graph
	                            option, the substep loop runs as one kernel per size class: each scene's
	                            workgroup builds its rows, solves, writes the impulses back and
	                            integrates its bodies, substep after substep, with no synchronisation
	                            between scenes.  0: one launch per stage and substep.  Default 0 (the
	                            fused kernel holds 2 waves per SIMD against the solve's 4, and is faster
	                            only on small single scenes).  Does not change results. */
	int sleeping;            /* 1: bodies fall asleep and wake as the reference's do (setSleepingEnabled,
	                            PhysicsEngine.cpp:1138-1140): movement history and sleep counters at the
	                            step start (:65-87, bodySleepy :1221-1254), an island whose bodies are all
	                            ready to sleep sleeps and is not solved (:1598-1606), an asleep body that
	                            a solve moves faster than 0.25 * sleep_vel_coeff * |g| wakes its island
	                            (:1124-1135), asleep bodies are not integrated and their contacts do not
	                            expire (PersistentConstraint.h:74).  Island modes only: a
	                            CP_PGS_COLORED step with sleeping on returns CP_ERR_STATE.  Default 0
	                            (the reference's default is on). */
	int sleep_tick_threshold; /* non_sleepy_tick_threshold (PhysicsEngine.h:220); default 3 */
	double sleep_vel_coeff;   /* vel_sleep_coeff (PhysicsEngine.h:222); default 0.1 */
	double sleep_ang_vel_eps; /* ang_vel_eps; default 0.1 */
	double sleep_accel_coeff; /* accel_sleep_coeff; default 0.044 */
	double sleep_delay;       /* sleep_delay (s) (PhysicsEngine.h:219); default 0.5.  The reference's
	                            setSleepParameters takes a sleep_assesment_time it never stores
	                            (PhysicsEngine.cpp:1186-1191), so its delay stays 0.5 */
} cp_config;

typedef struct cp_world cp_world;

void cp_config_default(cp_config* cfg);
const char* cp_status_string(int status);

/* World lifetime.  Replaces constructing a phyz::PhysicsEngine. */
int cp_world_create(const cp_config* cfg, cp_world** out);
int cp_world_destroy(cp_world* w);
int cp_world_set_config(cp_world* w, const cp_config* cfg);

/* Shapes (body-frame reference geometry, device resident).
 * cp_shape_add_hull_exact takes a polyhedron exactly as the reference stores
 * it after construction (Polyhedron ctor, ConvexPrimitive.cpp:582-658, and
 * computeGaussMap :861-896): faces with winding fixed, per-face
 * normal_calc_index / normalDirection (:783-812), Gauss normals, redundancy
 * flags, cached self-extrema (min_pid, max_pid, ref_idx) and (min, max,
 * ref_value), arcs and edges.  Used to replay reference states bit for bit. */
int cp_shape_add_hull_exact(cp_world* w, int n_pts, const double* pts, int n_faces, const int64_t* face_off,
                            const int64_t* face_verts, const int64_t* surface_id, const int64_t* normal_calc,
                            const int64_t* normal_dir, const double* gauss_normal, const int64_t* gauss_red,
                            const int64_t* gauss_ints, const double* gauss_vals, int n_arcs, const int64_t* arcs,
                            int n_edges, const int64_t* edges, int* shape_id);
/* Sphere (Sphere(center, radius), ConvexPrimitive.h:76-102): centre in the body
 * frame (the reference keeps the primitive's offset from the centre of mass). */
int cp_shape_add_sphere(cp_world* w, const double* center, double radius, int* shape_id);
/* Cylinder (Cylinder(center, radius, height, axis, 12, 8), ConvexPrimitive.cpp:222-256)
 * exactly as the reference stores its body-frame reference geometry:
 * params8 = centre xyz, height axis xyz, radius, height; gauss42 = 14 Gauss
 * vertices (cap normals, then the 12-gon ring); top24/bot24 = the 8-gon cap
 * polygons (radius r / cos(pi/8)). */
int cp_shape_add_cylinder(cp_world* w, const double* params8, const double* gauss42, const double* top24, const double* bot24,
                          int* shape_id);
/* Host-only: the reference geometry of Cylinder(center, radius, height) with
 * height axis +y, in the cp_shape_add_cylinder layout. */
int cp_cylinder_build(const double* center, double radius, double height, double* params8, double* gauss42, double* top24,
                      double* bot24);
/* Static triangle mesh (StaticMeshGeometry(const MeshInput&), Geometry.cpp:
 * 678-820): n_pts points (xyz, body frame), n_tris triangles as 3 point
 * indices each (counter-clockwise seen from the side the normal faces) and
 * per-triangle restitution, kinetic and static friction (TriIndices::material).
 * A body with this shape must be FIXED (createRigidBody(StaticMeshGeometry),
 * PhysicsEngine.cpp:498-513); its triangles and AABB tree are placed at its
 * pose when cp_world_set_state is called. */
int cp_shape_add_mesh(cp_world* w, int n_pts, const double* pts, int n_tris, const int64_t* tri_idx, const double* tri_mat3,
                      int* shape_id);
/* Convex polyhedron from points and face vertex lists (Polyhedron(points,
 * surface_vertex_indices), ConvexPrimitive.h:216): builds surfaces, edges and
 * the Gauss map on the host the way the reference constructor does.  Points
 * are given in the body frame (centre of mass at the origin). */
int cp_shape_add_polyhedron(cp_world* w, int n_pts, const double* pts, int n_faces, const int64_t* face_off,
                            const int64_t* face_verts, int* shape_id);
/* Host-only: the polyhedron cp_shape_add_polyhedron would build, returned in
 * the cp_shape_add_hull_exact layout.  Caller buffers: face_verts_out[total
 * face vertices] (winding fixed), per face surface_id/normal_calc/normal_dir/
 * gauss_red [n_faces], gauss_normal [3 n_faces], gauss_ints/gauss_vals
 * [3 n_faces]; arcs_out/edges_out [2 * cap_pairs] with counts returned. */
int cp_polyhedron_build(int n_pts, const double* pts, int n_faces, const int64_t* face_off, const int64_t* face_verts,
                        int64_t* face_verts_out, int64_t* surface_id, int64_t* normal_calc, int64_t* normal_dir,
                        double* gauss_normal, int64_t* gauss_red, int64_t* gauss_ints, double* gauss_vals, int cap_pairs,
                        int64_t* arcs_out, int* n_arcs, int64_t* edges_out, int* n_edges);
/* Mass properties of a closed polyhedron of uniform density:
 * out = mass, inertia tensor about the centre of mass [9], centre of mass [3]
 * (calculateMassProperties, RigidBody.cpp:427-640). */
int cp_polyhedron_mass_properties(int n_pts, const double* pts, int n_faces, const int64_t* face_off,
                                  const int64_t* face_verts, double density, double* out13);
/* Host-only: the polyhedron of a body created as createRigidBody(
 * ConvexUnionGeometry(Polyhedron(points, faces), material)) (RigidBody ctor,
 * RigidBody.cpp:13-38): pts_out[n_pts][3] = body-frame points (identity
 * rotation and -com applied to the points as given, recomputeFromReference,
 * ConvexPrimitive.cpp:715-729); mp19 = mass, reference tensor about the
 * centre of mass, its inverse (Mat3::inverse, Mat3.h:57-68); com3 = centre of
 * mass in the input frame.  The body's Gauss map, with the cached
 * self-extrema of the points as given, is what cp_polyhedron_build returns
 * for the same input. */
int cp_polyhedron_body(int n_pts, const double* pts, int n_faces, const int64_t* face_off, const int64_t* face_verts,
                       double density, double* pts_out, double* mp19, double* com3);

/* Bodies.  Replaces PhysicsEngine::createRigidBody (PhysicsEngine.cpp:486-496).
 * movement: 0 DYNAMIC, 1 FIXED (RigidBody.h:22).  mass_props: mass,
 * reference tensor[9], reference inverse tensor[9].  material: density,
 * restitution, kinetic and static friction (ConvexPrimitive.h:17-34). */
int cp_body_add(cp_world* w, int shape_id, int movement, const double* mass_props, const double* material, int scene,
                int* body_index);
/* Bulk form of cp_body_add for n bodies (arrays of n, n, 19n, 4n, n). */
int cp_body_add_many(cp_world* w, int n, const int* shape_ids, const int* movement, const double* mass_props,
                     const double* material, const int* scene, int* first_index);
int cp_world_body_count(cp_world* w, int* n);

/* State transfer (host <-> HBM).  inertia (nullable): world tensor[9] and
 * inverse[9] per body (RigidBody::tensor / invTensor); when null they are
 * recomputed from the orientation (RigidBody::updateInertiaTensor). */
int cp_world_set_state(cp_world* w, const double* state13, const double* inertia18);
int cp_world_get_state(cp_world* w, double* state13);
/* World inertia tensor[9] and its inverse[9] per body (RigidBody::tensor,
 * invTensor, RigidBody.h), the values cp_world_set_state takes back: state
 * and inertia together restore a body bit for bit. */
int cp_world_get_inertia(cp_world* w, double* inertia18);
int cp_world_get_aabbs(cp_world* w, double* aabb6);
int cp_world_set_contacts(cp_world* w, int n, const int64_t* ints6, const uint64_t* magic2, const double* f21);
int cp_world_get_contact_count(cp_world* w, int* n);
int cp_world_get_contacts(cp_world* w, int cap, int64_t* ints6, uint64_t* magic2, double* f21, int* n);

/* Island search root order for CP_PGS_REFERENCE_ORDER: body ids (index + 1)
 * in the order PhysicsEngine::sleepOrSolveIslands visits its
 * unordered_map (PhysicsEngine.cpp:1540).  cp_reference_node_order fills the
 * order libstdc++ produces for ids 1..n inserted in sequence. */
int cp_world_set_node_order(cp_world* w, int n, const int64_t* ids);
int cp_reference_node_order(int n, int64_t* ids);

/* Joint kinds (cp_joint_add's type) and their degrees (rows). */
typedef enum {
	CP_JOINT_DISTANCE = 1,      /* addDistanceConstraint, PhysicsEngine.cpp:623-658; 1 row */
	CP_JOINT_BALL_SOCKET = 3,   /* addBallSocketConstraint, :664-687; 3 rows */
	CP_JOINT_SLIDING_HINGE = 4, /* addSlidingHingeConstraint, :860-890; 4 rows */
	CP_JOINT_HINGE = 5,         /* addHingeConstraint, :693-725; 5 rows */
	CP_JOINT_WELD = 6,          /* addWeldConstraint, :896-920; 6 rows */
	CP_JOINT_SLIDER = 7,        /* addSliderConstraint, :773-806; 5 rows */
	CP_JOINT_CONE_LIMIT = 8,    /* addConeLimitConstraint, :926-945; 1 projected row while active */
	CP_JOINT_TWIST_LIMIT = 9,   /* addTwistLimitConstraint, :952-970; 1 projected row while active */
	CP_JOINT_MOTOR = 10,        /* addMotorConstraint, :725-760; 1 row while powered or outside its angle range */
	CP_JOINT_PISTON_DRIVE = 11, /* addPistonConstraint, :808-846: the piston's drive, 1 row while powered */
	CP_JOINT_SLIDE_LIMIT = 12,  /* addPistonConstraint: the piston's slide limit, 1 row while exceeded */
	CP_JOINT_SPRING = 13        /* addSpring, :974-995; no row: impulses at each substep's start */
} cp_joint_kind;

/* Powered modes of motors and piston drives (PoweredConstraitMode). */
typedef enum {
	CP_DRIVE_OFF = 0,
	CP_DRIVE_CONSTANT = 1,        /* constant torque / force: max */
	CP_DRIVE_TARGET_VELOCITY = 2, /* max torque / force, target velocity */
	CP_DRIVE_TARGET_POSITION = 3  /* max torque / force, target angle / position */
} cp_drive_mode;

/* Joints between bodies b1 and b2 of one scene; the pair stops colliding
 * (disallowCollision).  geo14 holds the joint as the reference stores it:
 * the tracked attach points in the body frames of b1 and b2
 * (RigidBody::trackPoint, RigidBody.cpp:134-137: local_coord_origin +
 * attach point), two body-frame axes - the hinge axes (Hinge::b1/
 * b2_rot_axis_body_space) or the normalised slide axes (Slider /
 * SlidingHinge ::b1/b2_slide_axis_body_space); for a distance joint the
 * first slot holds the target distance (Distance::target_distance, the
 * current distance when the reference is given -1); 0 otherwise - and the
 * position and rotation correction strengths.  The limits keep their angles
 * in the first attach-point slots (cone: the maximum angle; twist: the
 * minimum and maximum angles), the body-frame directions in the axis slots,
 * 0 as position strength and the rotation hardness as rotation strength.
 * A motor (on a hinge or sliding hinge: Motor takes the base joint's b1 and
 * b2 rotation axes and rotation hardness) holds its minimum and maximum
 * angles in the first attach-point slots and the two body-frame axes.  A
 * piston (on a slider or sliding hinge) is two joints, added in this order
 * (the order Piston::addSolverConstraints adds its rows): its slide limit
 * (12) and its drive (11), each with the base joint's attach points, b1's
 * slide axis and position strength; the slide limit holds the minimum and
 * maximum positions in the second axis slot.
 * A spring (13) holds its stiffness, damping and resting length in the first
 * axis slot (a negative length: the current distance, measured from the
 * bodies at the next step), adds no row and leaves the pair colliding; its
 * impulses are applied at the start of every substep, before any row is
 * built, in the constraint graph's edge order (PhysicsEngine.cpp:1313-1355).
 * Worlds with springs do not sleep: a step with sleeping on returns
 * CP_ERR_STATE.
 * Joints are solved in the island modes; a step of a world with joints in
 * CP_PGS_COLORED mode returns CP_ERR_STATE.  Kinds 1-7 are holonomic
 * constraints (isHolonomicConstraint) and join holonomic systems; kinds
 * 8-12 are inequality rows outside them.  Each substep a limit re-evaluates
 * its angle (ConeRotationLimit / TwistLimit::updateSolverConstraints,
 * PersistentConstraint.cpp:233-272) and, inside its range, clears its
 * impulse; a motor re-measures its angle and a piston its position (:300-329).
 * An island holds an inequality row when the joint was active at its
 * previous evaluation (a motor or drive also by its current powered mode). */
int cp_joint_add(cp_world* w, int type, int b1, int b2, const double* geo14, int* joint_index);
int cp_world_joint_count(cp_world* w, int* n);
/* Powered mode of motor or piston drive k: setMotorOff / ConstantTorque /
 * TargetVelocity / TargetPosition and setPistonOff / ConstantForce /
 * TargetVelocity / TargetPosition (PhysicsEngine.cpp:1683-1912).  Off keeps
 * the other parameters; the other modes set max and, for the targets,
 * target.  With sleeping on, both bodies receive a waking action.  A new
 * motor or drive is off. */
int cp_joint_set_drive(cp_world* w, int k, int mode, double max, double target);
/* Constraint force mixing of joint k's rows: setConstraintUseCustomCFM with
 * cfm >= 0, setConstraintUseGlobalCFM with NaN (PhysicsEngine.cpp:1012-1026;
 * CFM::getCFMValue, CFM.h:5-12).  The holonomic block factor keeps the
 * world's CFM, as the reference's computeInverse does (:1527-1531). */
int cp_joint_set_cfm(cp_world* w, int k, double cfm);
/* Joint solver state, 12 doubles per joint: accumulated impulse (6, unused
 * components 0), basis u and w of the current rows of hinges, sliders and
 * sliding hinges (0 for the others) (the PersistentConstraint ::constraint
 * objects, ConstraintSolver.h); a limit holds is_inactive (1 or 0) and, for a
 * twist limit, its tracked current_angle where u is (a new limit: 1, 0); a
 * motor its angular position and prev_velocity, a piston drive its position
 * and prev_velocity, a slide limit !slide_limit_exceeded and the position
 * (getMotorAngularPosition / getPistonPosition, PhysicsEngine.cpp:1914-1930),
 * a spring its resting length where u.x is.
 * A new motor or piston measures its position, and a spring given a negative
 * length its resting length, from the bodies at the next step or state read
 * (the Motor / Piston constructors, addSpring).
 * Setting it marks the holonomic systems for a rebuild of their rows and
 * factors at the next step. */
int cp_world_set_joint_state(cp_world* w, const double* state12);
int cp_world_get_joint_state(cp_world* w, double* state12);
/* Rows of joint k built by the last step: J_a, J_b [n][6], M^-1 J^T of each
 * body [6][n] (the constraints' constructors, ConstraintSolver.cpp:348-900);
 * the degree n (1 to 6) is returned. */
int cp_world_get_joint_rows(cp_world* w, int k, double* ja36, double* jb36, double* ita36, double* itb36, int* n);
/* Holonomic systems (HolonomicSystem, HolonomicBlockSolver.cpp), solved by
 * block LDL at the first substep when hol_iters > 0.  By default the world
 * finds them as maintainHolonomicSystems does (PhysicsEngine.cpp:1292-1307,
 * :1466-1512, :1746-1798) at the first step after joints are added;
 * cp_world_set_holonomic_systems sets them explicitly instead (joint indices
 * of system q at joints[off[q] .. off[q+1]), in constraint order) and turns
 * that search off. */
int cp_world_set_holonomic_systems(cp_world* w, int n_sys, const int64_t* off, const int64_t* joints);
int cp_world_holonomic_system_count(cp_world* w, int* n);
/* System k: constraint order og[n], elimination order[n] (joint indices),
 * block location table (lower triangle, column major, -1 empty), vector
 * offsets vloc[n] and the factor buffer; sizes4 = n, table size, degree,
 * factor size.  Any output pointer but sizes4 may be null. */
int cp_world_get_holonomic_system(cp_world* w, int k, int64_t* og, int64_t* order, int64_t* table, int64_t* vloc, double* factor,
                                  int64_t* sizes4);

/* One or more full steps.  Replaces PhysicsEngine::timeStep (PhysicsEngine.cpp:50-471). */
int cp_world_step(cp_world* w, int n_steps);

/* Stage entry points over the current state (parity seams of PhysicsEngine.cpp:91-335). */
/* Broadphase candidate set (AABBTree::getAllCollisionCandidates, AABB_Tree.h:153-224):
 * sorted (id_lo, id_hi) pairs. */
int cp_broadphase(cp_world* w, int cap, int64_t* pairs2, int* n);
/* Narrowphase of given (id_lo, id_hi) pairs: per pair the culled manifold
 * (detectCollision + cull_manifold, CollisionDetect.cpp:57-152, :2795-2830).
 * counts[n]; heads[n][4] = normal xyz, max depth; pts[n][4][4] = pos xyz, depth;
 * magic[n][4][2]. */
int cp_narrowphase(cp_world* w, int n, const int64_t* pairs2, int* counts, double* heads, double* pts, uint64_t* magic);
/* cp_narrowphase without allocation.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same inputs and
 * outputs as cp_narrowphase, computed by the same kernels on a fixed device
 * array of the library (about 6.4 MB per GPU, kept until the process ends)
 * instead of the world's growing buffers.  At most 16384 pairs per call
 * (CP_ERR_CAPACITY above); worlds holding a static mesh are refused
 * (CP_ERR_STATE: mesh pairs expand into many slots).  Every call on one GPU
 * uses that GPU's array, so calls that overlap on one GPU race. */
int cp_narrowphase_noalloc(cp_world* w, int n, const int64_t* pairs2, int* counts, double* heads, double* pts, uint64_t* magic);
/* Every manifold the engine adds for the given pairs, in pair order (a convex
 * body on a static mesh gets one per group of near-parallel triangle contacts,
 * PhysicsEngine.cpp:223-277): pair_index[cap] (index into pairs2), counts[cap],
 * heads[cap][4], pts[cap][4][4], magic[cap][4][2]; *n_out = manifolds found
 * (CP_ERR_CAPACITY when above cap). */
int cp_narrowphase_multi(cp_world* w, int n, const int64_t* pairs2, int cap, int64_t* pair_index, int* counts, double* heads,
                         double* pts, uint64_t* magic, int* n_out);
/* Sweep schedule of the last step: in CP_PGS_REFERENCE_ORDER mode the island
 * rows; in CP_PGS_ISLAND_COLORED mode the same islands with each island's rows
 * in colour order; in CP_PGS_COLORED mode (with diagnostics bit 1) one row
 * list per scene in colour order.  Rows are (contact index in the post-narrowphase cache,
 * kind 0 normal / 1 friction; or joint index, kind 2), with group offsets. */
int cp_world_last_islands(cp_world* w, int cap_rows, int64_t* rows2, int cap_offs, int64_t* offs, int* n_rows, int* n_islands);

/* With the coloured schedule recorded (CP_PGS_COLORED, diagnostics bit 1):
 * first[s] = index of scene s's first contact in the post-narrowphase cache
 * the schedule's rows index, first[n_scenes] = cache size; *n = n_scenes + 1.
 * Lets a caller replay one scene of a batched world on its own (the scene's
 * contacts are one contiguous range of that cache).  No counterpart in the
 * reference, which steps one scene per engine. */
int cp_world_last_scene_contacts(cp_world* w, int cap, int64_t* first, int* n);

/* Per-scene non-finite flags (the reference's NaN asserts, PhysicsEngine.cpp:
 * 1121-1122, compiled out in its Release build): flags[s] = 1 when a dynamic
 * body of scene s has held a NaN or infinite pose or velocity at the end of a
 * step since its state was last set (cp_world_set_state clears every flag,
 * admission clears the admitted bodies').  Scenes are the world's scenes in
 * order, or a serving world's body ranges in order.  Other scenes are not
 * affected: scenes never interact.  *n = number of scenes. */
int cp_world_scene_flags(cp_world* w, int cap, int* flags, int* n);

/* FP64 residual of the last substep's velocity solve (needs diagnostics bit 0):
 * out[0] = ||r|| over all live rows of the natural map r = lambda - proj(lambda +
 * A^-1 (target - J dv)) (contact: lambda >= 0; friction: disk of radius
 * mu * lambda_n), out[1] = ||lambda||, out[2] = out[0] / out[1], out[3] = rows. */
int cp_world_last_residual(cp_world* w, double* out4);

/* Serving world (continuous batching of independent scenes).  The nearest
 * reference interfaces are createRigidBody / removeRigidBody
 * (PhysicsEngine.cpp:486-513, :515-518) called on a running engine; scenes in
 * one serving world never interact.
 * cp_world_serving_init: turns an empty world (no bodies, no joints) into a
 *   serving world of `capacity` body slots (its HBM budget).
 * cp_world_admit_scene: stages a scene of n bodies into the first free slot
 *   range that holds it: shape_ids[n] (no meshes; at most 32 points and 64
 *   faces each), movement[n] (0 dynamic, 1 fixed), mp[n][19] (mass, I_ref,
 *   I_ref^-1), material[n][4], state13[n][13]; *handle names it.
 *   CP_ERR_CAPACITY when no free range is large enough (the caller queues).
 * cp_world_retire_scene: frees the scene's range and drops its contacts.
 * cp_world_serving_commit: applies staged admissions and retirements (the next
 *   cp_world_step does so itself).
 * cp_world_get_scene_state: the scene's bodies, state13[n][13]; with
 *   state13 == NULL only *n.
 * cp_world_serving_info: info4 = capacity, free slots, largest free range,
 *   admitted scenes. */
int cp_world_serving_init(cp_world* w, int capacity);
int cp_world_admit_scene(cp_world* w, int n, const int* shape_ids, const int* movement, const double* mp, const double* material,
                         const double* state13, int* handle);
int cp_world_retire_scene(cp_world* w, int handle);
int cp_world_serving_commit(cp_world* w);
int cp_world_get_scene_state(cp_world* w, int handle, double* state13, int* n);
/* The results of many admitted scenes in one call (one synchronisation, one
 * gather on the device, one copy back), the serving loop's per-tick read of
 * the scenes it retires: state13 receives their rows back to back in handle
 * order, counts[k] scene k's body count, status[k] CP_OK or CP_ERR_NONFINITE
 * (a body of that scene went non-finite).  CP_ERR_CAPACITY when the scenes
 * hold more than cap_rows rows.  No counterpart in the reference, which reads
 * its bodies one by one (RigidBody getters, RigidBody.h:25-68). */
int cp_world_get_scene_states(cp_world* w, int n, const int* handles, int cap_rows, double* state13, int* counts, int* status);
int cp_world_serving_info(cp_world* w, int* info4);

/* Snapshots and trajectories.  The reference records runs as frames of 7
 * doubles per body (position xyz, orientation r i j k; ImageDemo.h:98-133)
 * and resumes a step from the contact cache its constraint graph holds
 * (PhysicsEngine.cpp:1062-1111).
 *
 * cp_world_save_snapshot / cp_world_load_snapshot: the whole dynamic state of
 *   a world in one file, enough to resume it bit for bit in a world built
 *   with the same bodies (and joints): "CPSNAP2\n", int64 counts[4] =
 *   (bodies, contacts, joints, sleep), then state13[bodies], inertia18[bodies],
 *   contact rows (ints6, magic2, f21 per contact, the cp_world_get_contacts
 *   layout), joint state12[joints], powered modes [joints][4] (mode, max,
 *   target velocity, target position), then the solver objects a joint's
 *   update may leave standing: int64 row projection [joints], rows
 *   [joints][196] and pseudo impulses [6][joints]; when sleep = H + 1 > 0 (the world has
 *   had sleeping on), the sleep state: int64 [bodies][6] (asleep, asleep at
 *   the last contact removal, non_sleepy_tick_count, history length, history
 *   head, sleep disabled), sleep_ready_counter[bodies], and the movement
 *   history ring as doubles [H][6][bodies]; then a 64-bit FNV-1a checksum of
 *   every byte before it; all little-endian.  Loading checks the magic, the
 *   checksum and the body and joint counts (CP_ERR_IO otherwise) and sets
 *   state, inertia, contacts, joint state and solver objects, powered modes
 *   and sleep state (pending waking actions are not saved).
 * cp_world_trajectory_open: from the next cp_world_step on, appends one frame
 *   per step to `path` (created or truncated): per body, in index order,
 *   com xyz and orientation r i j k as doubles after the step.
 *   cp_world_trajectory_close ends the recording (also done by destroy).
 * cp_trajectory_frames / cp_trajectory_read (host-only): the number of
 *   frames of a file for n_bodies bodies, and frame `frame` as poses7[n][7]. */
int cp_world_save_snapshot(cp_world* w, const char* path);
int cp_world_load_snapshot(cp_world* w, const char* path);
int cp_world_trajectory_open(cp_world* w, const char* path);
int cp_world_trajectory_close(cp_world* w);
int cp_trajectory_frames(const char* path, int n_bodies, int* n_frames);
int cp_trajectory_read(const char* path, int n_bodies, int frame, double* poses7);

/* Serving: evicting a scene and admitting it again later, resuming exactly
 * where it stopped.
 * cp_world_scene_snapshot: scene `handle`'s bodies (state13[n], world
 *   inertia18[n]) and its contact rows in the cp_world_get_contacts layout,
 *   with body ids relative to the scene (1..n; MagicID primitive halves
 *   shifted with them); with state13 == NULL only *n_bodies and *n_contacts.
 *   CP_ERR_CAPACITY when the scene has more than cap_contacts contacts.
 * cp_world_admit_scene_resume: cp_world_admit_scene with that snapshot's
 *   inertia and contact rows applied when the admission is committed; the
 *   scene then steps as if it had never left, wherever its new range is. */
int cp_world_scene_snapshot(cp_world* w, int handle, int cap_contacts, double* state13, double* inertia18, int64_t* ints6,
                            uint64_t* magic2, double* f21, int* n_bodies, int* n_contacts);
int cp_world_admit_scene_resume(cp_world* w, int n, const int* shape_ids, const int* movement, const double* mp,
                                const double* material, const double* state13, const double* inertia18, int n_contacts,
                                const int64_t* ints6, const uint64_t* magic2, const double* f21, int* handle);

/* Sleeping (cp_config.sleeping).
 * cp_world_get_sleep_state: per body asleep (RigidBody::getAsleep), sleep_ready_counter,
 *   non_sleepy_tick_count and the length of its movement history, as out4[n][4].
 * cp_world_wake_bodies: RigidBody::alertWakingAction (RigidBody.cpp:340-345) for the listed
 *   bodies (what the reference's setters setVel, setPos, setOrientation, ... call): their
 *   counters and histories reset now, and at the next step start each one's island wakes
 *   (PhysicsEngine.cpp:66-69) - the island of the constraint graph as it stands then.
 * cp_world_set_sleep_disabled: RigidBody::setSleepDisabled for the listed bodies (a body with
 *   sleep disabled is never sleepy, so its island never sleeps).  The reference leaves the
 *   flag uninitialised in its constructors (RigidBody.h:89); here it starts at 0. */
int cp_world_get_sleep_state(cp_world* w, double* out4);
int cp_world_wake_bodies(cp_world* w, int n, const int* idx);
int cp_world_set_sleep_disabled(cp_world* w, int n, const int* idx, int disabled);

/* Batched raycasts (PhysicsEngine::raycastFirstIntersection, PhysicsEngine.cpp:
 * 582-609, for n rays at once) against the world geometry of the last step:
 * ray k starts at origins3[k] with unit direction dirs3[k] and sees the
 * bodies of scene scene_of_ray[k] (-1, or scene_of_ray NULL: every body;
 * scenes never see each other).  hit_body[k] = index of the first body hit
 * (least distance t >= 0; -1 none), hit7[k] = hit point xyz, surface normal
 * xyz, t (zeros when nothing is hit).  Bodies are tested as
 * RigidBody::checkRayIntersection tests them (RigidBody.cpp:185-204), boxes
 * first (AABB::rayIntersectsAABB, AABB.h:143-173).  A tie in t between two
 * bodies goes to the lower index; the reference takes the first candidate
 * of its body tree's traversal.  The ignore list of the reference query is
 * not taken. */
int cp_world_raycast(cp_world* w, int n, const double* origins3, const double* dirs3, const int* scene_of_ray, int* hit_body,
                     double* hit7);

/* Timing of the last cp_world_step call, per phase, in milliseconds (HIP events):
 * [0] gravity + broadphase, [1] narrowphase, [2] persistence + schedule,
 * [3] substeps (rows + PGS + integrate), [4] geometry refresh,
 * [5] PGS kernels only (sum over substeps), [6] PGS launches, [7] live contacts.
 * With the fused substep kernels (cp_config.fused) the solve is not timed on
 * its own: [5] = [3] and [6] = 0. */
int cp_world_last_timing(cp_world* w, double* ms8);

#ifdef __cplusplus
}
#endif
#endif
