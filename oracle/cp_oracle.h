// cp_oracle — FP64 CPU restatement of ConstraintPhysics' per-step contact
// pipeline.  TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and the
// bench's cpu_baseline leg may load it, and only as the checker.  The product
// (constraintphysics_amd) never links or calls it.
//
// Every function follows the reference's operation order so that, fed the same
// per-step state, it reproduces the reference's FP64 results bit for bit (the
// reference is compiled with -ffp-contract=off; so is this file).  Each
// function cites the reference file:line it restates.  Parity of each part is
// pinned against golden vectors dumped from the compiled reference
// (tests/golden/, produced by oracle/ref_harness.cpp) — see tests/test_oracle.py.
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_world orc_world;

// Engine parameters (PhysicsEngine.h:174-216 defaults unless set).
typedef struct {
	double dt;            // step time
	double gravity[3];
	int vel_iters, pos_iters, substeps;
	double cfm;           // global constraint-force mixing
	double contact_hardness;
	int contact_life;
	int drop_nonfinite;   // 1: drop non-finite manifold points (the product's defect fix); 0: as the reference
	int hol_iters;        // holonomic block-solver iterations (setPGSIterations' third argument); 0 = none
} orc_config;

orc_world* orc_create(const orc_config* cfg);
void orc_destroy(orc_world* w);

// Add a body.  movement: 0 dynamic, 1 fixed.  mass_props: mass, I_ref[9], I_ref_inv[9].
// material: density, restitution, kinetic mu, static mu.
// Polyhedron reference geometry in the body frame: n_pts points, n_faces faces
// (face_off[n_faces+1] into face_verts), per-face normal_calc_index and
// normal_direction (ConvexPrimitive.cpp:791-812), Gauss map (normals, redundant
// flags, cached extrema ints (min_pid, max_pid, ref_idx) and vals (min, max,
// ref_val)), arcs, edges.  Returns body index (ids are index+1).
int orc_add_hull_body(orc_world* w, int movement, const double* mass_props, const double* material,
                      int n_pts, const double* pts, int n_faces, const int64_t* face_off, const int64_t* face_verts,
                      const int64_t* surface_id, const int64_t* normal_calc, const int64_t* normal_dir,
                      const double* gauss_normal, const int64_t* gauss_red, const int64_t* gauss_ints, const double* gauss_vals,
                      int n_arcs, const int64_t* arcs, int n_edges, const int64_t* edges);

// Sphere body: centre in the body frame and radius (Sphere, ConvexPrimitive.h:76-102).
int orc_add_sphere_body(orc_world* w, int movement, const double* mass_props, const double* material, const double* center,
                        double radius);

// Cylinder body (Cylinder, ConvexPrimitive.h:159-210) as the reference stores its
// body-frame reference: params8 = centre xyz, height axis xyz, radius, height;
// gauss42 = 14 Gauss vertices; top24/bot24 = the 8-gon cap polygons.
int orc_add_cylinder_body(orc_world* w, int movement, const double* mass_props, const double* material, const double* params8,
                          const double* gauss42, const double* top24, const double* bot24);

// Static triangle mesh body (RigidBody(StaticMeshGeometry), RigidBody.cpp:40-58,
// built by StaticMeshGeometry(const MeshInput&), Geometry.cpp:678-820): points,
// triangles (3 point indices each, counter-clockwise seen from the front) and
// per-triangle restitution, kinetic and static friction.  FIXED, centre of mass
// at the origin, identity orientation.  Returns the body index.
int orc_add_mesh_body(orc_world* w, int n_pts, const double* pts, int n_tris, const int64_t* tri_idx, const double* tri_mat3);
// Triangles of mesh body `body` whose boxes meet aabb6, in the order of
// AABBTree::getCollisionCandidatesWith (AABB_Tree.h:226-247); returns the count.
int orc_mesh_query(const orc_world* w, int body, const double* aabb6, int64_t* out, int cap);
// World triangle t of a mesh body: f40 = normal, 3 vertices, 3 edge out
// directions, up to 3 Gauss-region vectors, aabb 6, restitution, mu_k, mu_s, 0;
// i8 = 3 vertex ids, 3 edge ids, Gauss-region size, concave neighbour count.
void orc_mesh_triangle(const orc_world* w, int body, int t, double* f40, int64_t* i8);
// Narrowphase of one body pair returning every manifold the engine adds
// (several for a convex body on a mesh); returns the manifold count, up to cap
// written: heads[cap][4], counts[cap], pts / magic rows in manifold order.
int orc_narrowphase_multi(const orc_world* w, int id_lo, int id_hi, int cap, double* heads, int64_t* counts, double* pts,
                          uint64_t* magic);

// Joints (PhysicsEngine::addBallSocketConstraint / addHingeConstraint,
// PhysicsEngine.cpp:664-720).  type 3 ball-socket, 5 hinge; b1/b2 body
// indices; geo14 = tracked point in b1's frame (3), in b2's frame (3), hinge
// axis in b1's frame (3), in b2's frame (3), position and rotation correction
// strengths.  The pair stops colliding.  Returns the joint index.
int orc_add_joint(orc_world* w, int type, int b1, int b2, const double* geo14);
// Joint solver state rows of 11 doubles: impulse (5, zero padded), hinge u, w.
/* Custom CFM of joint k (setConstraintUseCustomCFM); NaN: the global CFM. */
void orc_set_joint_cfm(orc_world* w, int k, double cfm);
/* Island roots (body ids) for steps given no node order. */
void orc_set_node_order(orc_world* w, int n, const int64_t* ids);
/* Joint state as the reference's constructors leave it, from the current bodies. */
void orc_init_joint_state(orc_world* w);
/* Powered mode of a motor (10) or piston drive (11): 0 off, 1 constant
 * torque / force (max), 2 target velocity (max, target), 3 target position
 * (max, target).  Returns -1 for another kind. */
int orc_set_drive(orc_world* w, int k, int mode, double max, double target);
void orc_set_joint_state(orc_world* w, const double* s11);
void orc_get_joint_state(const orc_world* w, double* s11);
// Holonomic systems: n_sys systems, system k holds joints og[off[k]..off[k+1])
// in the order the reference's revaluateSystem collects them.  Rows and factors
// are built at the next step start.
void orc_set_hsys(orc_world* w, int n_sys, const int64_t* off, const int64_t* og);
int orc_hsys_count(const orc_world* w);
// Joint k's current rows: J_a, J_b [n][6] and M_a^-1 J_a^T, M_b^-1 J_b^T [6][n]; returns n.
int orc_get_joint_rows(const orc_world* w, int k, double* ja, double* jb, double* ita, double* itb);
// Builds pending systems' rows and factors now (as the next step would).
void orc_hsys_prepare(orc_world* w);
// System k's elimination order, block table, vector offsets and factor buffer;
// sizes4 = {n joints, table entries, degree, buffer doubles}.  Null arrays
// return sizes only.
void orc_get_hsys(const orc_world* w, int k, int64_t* order, int64_t* table, int64_t* vloc, double* buf, int64_t* sizes4);

// Body state rows of 13 doubles: com 3, q (r,i,j,k), vel 3, ang_vel 3.
void orc_set_state(orc_world* w, const double* state);
void orc_get_state(const orc_world* w, double* state);
// World inertia rows of 18 doubles: tensor 9, inverse tensor 9 (RigidBody.h:116-117).
void orc_set_inertia(orc_world* w, const double* inertia);
// Recompute world geometry/AABBs/inertia from the current state (RigidBody::updateGeometry
// and updateInertiaTensor, RigidBody.cpp:270-312) without changing the state.
void orc_refresh(orc_world* w);

// Contact cache rows: ints (b1, b2, memory_life, is_live, static_ready, pos_in_edge),
// magic (cid, bid), 21 doubles (local_b1 3, local_b2 3, normal_local_b1 3, pen, bounce,
// mu, imp_n, imp_t0, imp_t1, u 3, w 3).  Rows are grouped per edge in edge order.
void orc_set_cache(orc_world* w, int n, const int64_t* ints, const uint64_t* magic, const double* f);
int orc_cache_size(const orc_world* w);
void orc_get_cache(const orc_world* w, int64_t* ints, uint64_t* magic, double* f);

// Broadphase: all (lo, hi) body-id pairs with overlapping closed AABBs, sorted.
int orc_broadphase(const orc_world* w, int64_t* pairs, int cap);
void orc_get_aabbs(const orc_world* w, double* aabbs);
void orc_get_world_points(const orc_world* w, int body, double* pts);
// First body hit by the ray o3 + t d3 (t >= 0) among bodies [b0, b1) at the
// current world geometry (PhysicsEngine::raycastFirstIntersection,
// PhysicsEngine.cpp:582-609): body index or -1; out7 = hit point, surface
// normal, distance t.
int orc_raycast(const orc_world* w, const double* o3, const double* d3, int b0, int b1, double* out7);

// Narrowphase of one body pair (ids lo < hi): culled manifold as the engine
// would add it (PhysicsEngine.cpp:167-277).  Returns number of points (0 if no
// collision); head = normal 3, max_pen; pts rows = pos 3, pen; magic rows = cid, bid.
int orc_narrowphase(const orc_world* w, int id_lo, int id_hi, double* head, double* pts, uint64_t* magic);

// One full step (PhysicsEngine::timeStep without collision actions).
// node_order: body ids in the order the island search visits roots
// (PhysicsEngine.cpp:1540); NULL = ascending id.
void orc_step(orc_world* w, const int64_t* node_order);

// One full step with an explicit sweep schedule: n_groups row lists
// (rows[2*r] = index into the post-narrowphase cache, rows[2*r+1] = 0 normal /
// 1 friction, 2 joint; offs[n_groups+1]), solved as PGS_solve solves one
// island each.  With sleeping on, the groups must be the step's islands
// (their sleep decisions and wake lists come from the island search; a
// group is matched to the island of its first dynamic body); returns -1 when
// they are not, else 0.
int orc_step_scheduled(orc_world* w, int n_groups, const int64_t* rows, const int64_t* offs);

// Sleeping (PhysicsEngine::setSleepingEnabled / setSleepParameters,
// PhysicsEngine.cpp:1138-1190; defaults PhysicsEngine.h:218-224): enabled,
// params4 = sleep_delay, vel_sleep_coeff, ang_vel_eps, accel_sleep_coeff
// (NULL keeps them), tick_threshold = non_sleepy_tick_threshold (<= 0 keeps).
void orc_set_sleeping(orc_world* w, int enabled, const double* params4, int tick_threshold);
// Per body: asleep, sleep_ready_counter, non_sleepy_tick_count, history length.
void orc_get_sleep_state(const orc_world* w, double* out4);
// RigidBody::alertWakingAction (RigidBody.cpp:340-345) for the listed bodies:
// their islands wake at the next step start.
void orc_wake_bodies(orc_world* w, int n, const int64_t* idx);
// RigidBody::setSleepDisabled for the listed bodies.
void orc_set_sleep_disabled(orc_world* w, int n, const int64_t* idx, int disabled);

// Island rows of the last step: (cache index into the post-narrowphase cache, kind 0 normal / 1 friction).
int orc_last_islands(const orc_world* w, int64_t* rows, int64_t* offs, int cap);
int orc_last_mid_cache_size(const orc_world* w);
void orc_last_mid_cache(const orc_world* w, int64_t* ints, uint64_t* magic, double* f);

#ifdef __cplusplus
}
#endif
