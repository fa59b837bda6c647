// Device-side data layout of a world (one or many independent scenes) and the
// kernel launchers shared by the translation units of libcp_hip.so.
//
// Layout (all in HBM, structure-of-arrays, component-major so that a thread
// per body/contact reads coalesced):
//   bodies   com[3][nb] q[4][nb] vel/angv/pvel/pang[3][nb] (FP64), inertia
//            I/Iinv world and reference [9][nb], mass, movement, shape, scene.
//   shapes   body-frame hull: points, faces (vertex lists), Gauss map (face
//            normals, redundancy flags, cached self-extrema), Gauss arcs,
//            edges, per-vertex face/edge adjacency.  One shape may be shared.
//   world    per-body world hull points and world Gauss normals, AABBs.
//   pairs    broadphase output: u64 key lo<<32|hi (global body indices).
//   contacts the persistent cache, sorted by (pair key, insertion order), with
//            per-substep solver rows stored beside it.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cp_math.h"
#include "../../include/cp_abi.h"

namespace cpx {

// Workgroups of a dispatch go to the eight XCDs round-robin (block b to XCD
// b mod 8), and each XCD has its own L2.  xcd_block() renumbers the blocks so
// that each XCD runs one contiguous range of logical blocks; work items next
// to each other (a scene's pairs or contacts, which share bodies) then share
// an L2.  The grid must be a multiple of 8 (xcd_grid rounds it up).
__device__ __forceinline__ int xcd_block() {
	const int per = gridDim.x >> 3;
	return (blockIdx.x & 7) * per + (blockIdx.x >> 3);
}
static inline int xcd_grid(int blocks) { return (blocks + 7) & ~7; }

enum ShapeType { SHAPE_HULL = 0, SHAPE_SPHERE = 1, SHAPE_CYLINDER = 2, SHAPE_MESH = 3 };

struct Shapes {
	int n;
	int* type;
	int *pt_off, *face_off, *arc_off, *edge_off;  // [n+1] into the flat arrays below
	double* ref_pts;                              // [3*P] xyz interleaved
	int *fv_off, *fv;                             // per global face: vertex list offsets [F+1], vertices
	int *f_ncalc, *f_ndir, *f_sid;                // Surface normal_calc_index / direction / surface id
	double* g_norm;                               // [3*F] body-frame Gauss normals
	int *g_red, *g_minpid, *g_maxpid, *g_refidx;  // redundancy + cached extrema (per face)
	double *g_min, *g_max, *g_refval;
	int* arcs;                                    // [2*A] face-index pairs (shape-local)
	int* edges;                                   // [2*E] vertex-index pairs (shape-local)
	int *va_face_off, *va_faces;                  // per global vertex: adjacent faces (shape-local ids)
	int *va_edge_off, *va_edges;                  // per global vertex: adjacent edges (shape-local ids)
	double* radius;                               // sphere / cylinder radius
	double* height;                               // cylinder height
};

struct Bodies {
	int n;   // bodies (the stride of every per-body array)
	int nl;  // bodies the per-step kernels visit: all, or a serving world's occupied prefix
	double *com, *q, *vel, *angv, *pvel, *pang, *prev_com, *prev_q;
	double *I, *Iinv, *Iref, *Irefinv;
	double* mass;
	int *movement, *shape, *scene;
	double* mat;      // [3][n]: restitution, mu_kinetic, mu_static
	double* aabb;     // [6][n]: min xyz, max xyz
	int* wpt_off;     // [n+1] world point offsets
	double* wpts;     // [3*WP] world points (xyz interleaved)
	int* wgv_off;     // [n+1] world Gauss normal offsets
	double* wnorm;    // [3*WG]
	double* dv;       // [12][n] PGS velocity / pseudo-velocity deltas
	float* iinv32;    // [n][8] world inverse inertia in FP32 for the split solve: xx, xy, xz, yy, yz, zz (0 unless dynamic), 2 unused
	int4* geo;        // [n][3] per body, rebuilt at each narrowphase: (shape type, shape, world point offset, world normal
	                  // offset), (first point, points, first face, faces), (first arc, arcs, first edge, edges)
	int* color_mask;  // scratch for colouring (2 x u64 per body, as 4 ints)
	// sleeping (null while no body can be asleep): RigidBody::asleep, and the
	// asleep flags of the last step's first substep, when the reference
	// removes expired contacts (Contact::isExpired spares asleep bodies')
	int* asleep;
	const int* asleep_rm;
};

struct Params {
	double dt, sub_dt, gx, gy, gz, cfm, hardness, cutoff_vel;
	double cos_tol, sin_tol;
	double cull_c[4], cull_s[4];  // cos/sin(2*pi*i/4), host libm values
	int vel_iters, pos_iters, substeps, contact_life;
	int gyro_ticks;
	int defect_fixes;
	int precision;  // cp_precision (CP_PREC_FP64 or CP_PREC_SPLIT)
	int hol_iters;  // holonomic block-solver iterations (first substep only)
};

// Joints, SoA over n joints.  kind: 1 distance, 3 ball-socket, 4 sliding
// hinge, 5 hinge, 6 weld, 7 slider, 8 cone limit, 9 twist limit, 10 motor,
// 11 piston drive, 12 piston slide limit, 13 spring; type: the joint's degree
// (rows: 1, 3, 4, 5, 6, 5, then 1 for each inequality kind, 0 for a spring).  State arrays are
// component-major ([CP_JMAX][n], [3][n]).
#define CP_JMAX 6
#define CP_JSTATE 12  // joint state per joint at the ABI: CP_JMAX impulses, u, w
struct Joints {
	int n;
	int *type, *kind, *a, *b, *sys;  // sys: holonomic system, -1 none
	double *tp1, *tp2, *ax1, *ax2;  // [3][n] body-frame attach points and axes (distance: ax1 x = target distance)
	double *ps, *rs;                // correction strengths
	double *imp, *pimp, *himp;      // [CP_JMAX][n]
	double *u, *w;                  // [3][n] basis of the current rows (hinge, slider, sliding hinge);
	                                // cone / twist / slide limits: u x = inactive (1.0 / 0.0); twist: u y = tracked angle;
	                                // slide limit: u y = piston position; motor, piston drive: u x = position, u y = prev_velocity
	double* drive;                  // [4][n] powered mode (0 off, 1 constant, 2 target velocity, 3 target position), max, target velocity, target position
	                                // spring: ax1 = stiffness, damping, resting length as given; u x = resting length
	int* proj;                      // the rows' impulse projection (CP_PROJ_*)
	double* row;                    // CP_JRS doubles per joint
	double* cfm;                    // [n] custom CFM (setConstraintUseCustomCFM), NaN: the world's
};
// Kinds 8-12 are one-row inequality constraints, present in an island only
// while active and never part of a holonomic system: cone (8) and twist (9)
// limits (PersistentConstraint.h:145-215), motor (10), and a piston's drive
// (11) and slide limit (12) (:218-290).  A spring (13) is not holonomic either
// and adds no row.
__host__ __device__ __forceinline__ bool joint_is_inequality(int kind) { return kind >= 8; }
// Projection of a row's accumulated impulse (projectValidImpulse), low bits;
// CP_PROJ_NOPOS marks a row that the position pass skips (needsPosCorrect).
enum { CP_PROJ_NONE = 0, CP_PROJ_LOWER = 1, CP_PROJ_UPPER = 2, CP_PROJ_CLAMP = 3, CP_PROJ_NOPOS = 4 };
// Whether joint k adds its row to an island: limits while active, a slide
// limit while exceeded (flags of their last update), a motor while powered or
// outside its angle range (Motor::constraintIsActive, PersistentConstraint.cpp:
// 110-112), a piston drive while powered (Piston::pistonIsActive, :195-197).
struct JointAct {
	int n;
	const int* kind;
	const double *u, *drive, *tp1;
	__device__ __forceinline__ bool operator()(int k) const {
		const int kd = kind[k];
		if (kd < 8) return true;
		if (kd == 13) return false;  // springs add no row (Spring::isSolverConstraint)
		if (kd == 10) {
			double pos = u[k];
			return (drive[k] != 0.0 && drive[n + k] != 0) || pos < tp1[k] || pos > tp1[n + k];
		}
		if (kd == 11) return drive[k] != 0.0 && drive[n + k] != 0;
		return u[k] == 0.0;
	}
};
#define CP_J_BOUND 192  // a powered row's impulse bound (max_torque_impulse, max_impulse)
#define CP_J_TV 193     // a powered row's target velocity (real_target_velocity)
#define CP_JRS 196
enum JointRowField {
	CP_J_JA = 0,      // J_a [CP_JMAX][6]
	CP_J_JB = 36,     // J_b [CP_JMAX][6]
	CP_J_ITA = 72,    // M_a^-1 J_a^T [6][CP_JMAX]
	CP_J_ITB = 108,   // M_b^-1 J_b^T [6][CP_JMAX]
	CP_J_INV = 144,   // (J M^-1 J^T (1+cfm))^-1 [n][n]
	CP_J_TGT = 180,   // velocity targets [CP_JMAX]
	CP_J_PTGT = 186   // pseudo-velocity targets [CP_JMAX]
};

// Holonomic systems: per system the elimination order (joint indices), the
// block location table (lower triangle, column major, -1 = empty block),
// vector offsets, and slices of the factor buffer and its column scratch.
struct HSysDev {
	int nsys;
	int *nc, *ord_off, *order, *vloc, *tab_off, *table, *degree, *buf_off, *dbuf_off;
	double *buf, *dbuf;
	double* scratch;  // 3 * degree per system (at dbuf_off / 2): solve vectors
	// per system s and column c (in elimination order): the non-empty blocks
	// below the diagonal, rows ascending, as (cs_row, cs_loc)[cs_ptr[cp_off[s]
	// + c], cs_ptr[cp_off[s] + c + 1]); the diagonal block at dg_loc[ord_off[s] + c]
	int *cp_off, *cs_ptr, *cs_row, *cs_loc, *dg_loc;
	int lds_doubles;  // the largest system's block buffer plus column scratch (the wave factor stages them in LDS)
};

// enabled: the arrays are valid and asleep bodies are honoured (some body
// may be asleep); sleeping: islands may fall asleep this step.
struct SleepDev {
	int enabled, sleeping;
	int* asleep;        // [n]
	double* counter;    // [n] sleep_ready_counter
	int* ticks;         // [n] non_sleepy_tick_count
	int* hlen;          // [n] history entries held
	int* hhead;         // [n] ring position of the oldest entry
	double* hist;       // [hcap][6][n] (vel xyz, ang_vel xyz) per entry
	const int* proc;    // [n] or null: whether the step-start loop treats the body as awake (set
	                    // when waking actions changed asleep flags during that loop), else !asleep
	const int* disabled;  // [n] RigidBody::sleep_disabled
	int hcap;           // history length kept: (int)(sleep_delay / step_time)
	double delay, step_time, vel_eps, ang_vel_eps, accel_eps, ang_accel_eps, wake_vel, hist_need;
	int tick_threshold;
};

// Island schedule built on the device each step.  rows: (index, kind)
// pairs, kind 0 contact normal, 1 friction, 2 joint; an island's rows are
// [off[i], off[i+1]).  Its holonomic systems are sys[sys_off[i], sys_off[i] +
// sys_cnt[i]).  Coloured islands: an edge's rows are [eoff[e], eoff[e+1]),
// a colour's edges [coff[c], coff[c+1]), an island's colours [icoff[i],
// icoff[i] + ncol[i]), the last one a serial bucket when serial[i].
// jlin[2k], jlin[2k+1]: columns of joint k's bodies in the linearization-error
// vectors (stride nl): the body index for a dynamic body, a column of the
// island beyond the bodies for the others.  The island's bodies (every body
// of its component, in walk order) are vlist[voff[i], voff[i] + vcnt[i]).
struct IslandSched {
	int n, nl;
	const int *rows, *off, *sys, *sys_off, *sys_cnt, *jlin;
	const int *eoff, *coff, *icoff, *ncol, *serial;
	// coloured islands with island-local velocity deltas in LDS: the island's
	// dynamic bodies with rows bod[boff[i], boff[i] + nbod[i]) (local index =
	// position), contact k's bodies as local indices cloc[2k], cloc[2k+1],
	// joint k's jloc[2k], jloc[2k+1] (the zero column nbod[i] for non-dynamic
	// bodies); islands with more than lds_bodies - 1 bodies use b.dv.  cloc
	// null: none.
	const int *bod, *boff, *nbod, *cloc, *jloc;
	int lds_bodies;
	const int *vlist, *voff, *vcnt;
	SleepDev sl;  // sleeping: the wake check after each island solve
};

// Per-component counts of the island search (indexed by the component's
// root body), summed by an exclusive scan over node-map positions.
struct IslCnt {
	int n, rows, edges, bodies, adj, joints, notready;
};
struct IslCntSum {
	__host__ __device__ IslCnt operator()(const IslCnt& a, const IslCnt& b) const {
		return IslCnt{a.n + b.n, a.rows + b.rows, a.edges + b.edges, a.bodies + b.bodies, a.adj + b.adj, a.joints + b.joints,
		              a.notready | b.notready};
	}
};
// Island ranges: root body, row offsets (n + 1), walked-edge, body, stack
// and joint offsets, bodies per island.
struct IslTables {
	int *root, *off, *edge_off, *body_off, *body_cnt, *stack_off, *joint_off;
};
// Sleeping state (PhysicsEngine.h:218-224; RigidBody asleep,
// sleep_ready_counter, non_sleepy_tick_count, history).
// What the per-island walk reads and writes.
struct IslWalk {
	IslTables T;
	int nb, hol;
	const int *movement, *asleep, *aoff, *adjv, *me, *jord, *live, *ja, *jb, *jsys;
	JointAct jact;
	const uint64_t *adjk, *ckey;
	int *stack, *vlist, *elist, *vis, *rows, *ecol, *coff, *eorder, *icoff, *ncol, *serial, *eoff, *bod, *nbod, *loc, *cloc, *jloc;
	int *sys_seen, *sys, *slots, *jlin, *sys_cnt;
	uint64_t* col_used;
};

// Two side streams forked from and joined back into a main stream with events
// (independent launches run beside each other).
struct Fork {
	hipStream_t side[2];
	hipEvent_t fork, join[2];
};

// Narrowphase output, one slot per candidate pair.
struct Manifolds {
	int* count;       // [cap] points in the culled manifold (0 = none)
	double* normal;   // [3][cap]
	double* maxpen;   // [cap]
	double* pts;      // [cap*4*4]: pos xyz, pen
	uint64_t* magic;  // [cap*4*2]: cid, bid
	double* mat;      // [cap*4][3]: restitution, mu_k, mu_s of each point
};

// Static triangle meshes in world coordinates, one instance per mesh body
// (records as laid out by place_mesh: triangle MT_REC doubles / MT_IREC ints,
// tree nodes box[6] + (left, right), a leaf being (-1, triangle)); global
// triangle and node indices, per instance the root node.
struct MeshDev {
	int n_inst;
	const int* inst_of_body;  // [nb], -1 for other bodies
	const int* root;          // [n_inst]
	const double* tri;        // [T][MT_REC]
	const int* tri_i;         // [T][MT_IREC]
	const double* node_box;   // [N][6]
	const int* node_lr;       // [N][2]
};
#define MT_REC 40
#define MT_IREC 8
#define MESH_SLOTS 16   // manifold slots per convex-body / mesh pair
#define MESH_STACK 64   // tree traversal stack (depth + 1 entries suffice)

// Per (mesh pair, candidate triangle) manifold before merging.
#define MESH_PTS 40     // points of one triangle manifold (the clip pool bound)
struct MeshItems {
	int* slot;        // [C] first manifold slot of the pair
	int* tri;         // [C] global triangle index
	int* count;       // [C] points (0 = no manifold)
	double* head;     // [C][4] normal, max depth
	double* pts;      // [C][MESH_PTS][4]
	uint64_t* magic;  // [C][MESH_PTS][2]
	double* mat;      // [C][3]
	int* owner;       // [C] merge scratch
	uint64_t* avail;  // [C] cull scratch
};

// Contact cache (SoA).  Row data is rebuilt each substep for live contacts.
struct Contacts {
	int cap;
	uint64_t *key, *cid, *bid, *seq;
	double *lp1, *lp2, *ln1;  // [3][cap]
	double *pen, *bounce, *mu;
	int *life, *live, *sready;
	double *imp_n, *imp_t;  // imp_t [2][cap]; persistent warm-start impulses
	double *fu, *fw;        // [3][cap] friction basis of the last row build
	// Solver rows, one contiguous record of CP_RS doubles per contact (the
	// coloured sweep visits contacts out of memory order, so a record keeps a
	// row update within a few cache lines).  Layout: CP_R_* offsets below.
	double* row;
};

// A cached contact the reference removed at the previous step's first
// substep (removeExpiredContactConstraints, PhysicsEngine.cpp:1434-1446):
// memory_life was below 0 there (below -1 after the end-of-step decrement)
// and neither body was asleep then (Contact::isExpired,
// PersistentConstraint.h:74; prot: the asleep flags of that moment, null
// when no body sleeps).
__device__ __forceinline__ bool contact_removed(const Contacts& C, int k, const int* prot) {
	if (C.life[k] >= -1) return false;
	if (!prot) return true;
	return !prot[(int)(C.key[k] >> 32)] && !prot[(int)(C.key[k] & 0xFFFFFFFFu)];
}

// Record of CP_RS values (FP64: 512 B = 4 lines of 128 B; FP32: 256 B = 2
// lines).  The fields a sweep writes (impulses, static flag) share the first
// line.  Only the independent coefficients are stored: the linear parts of
// the Jacobians are -n | n (normal row) and u | -u, w | -w (friction rows),
// and the linear parts of M^-1 J^T are the inverse masses times those, so a
// row step rebuilds them from the basis block as 0 + (1/m) j, the value the
// full 6x6 product M^-1 J^T gives (its other five terms are signed zeros).
// Blocks start at multiples of 4 values and records are 128-byte aligned, so
// the blocks load as 16-byte vectors.
//
// The split solve stores a compact FP32 record of CP_RSF values (128 B, one
// cache line): the first 16 values as below with the inverse mass of body a
// in the padding slot, then the basis and the two lever arms (CP_F_*).  A row
// step rebuilds the angular Jacobians as cross products of the lever arms
// with the basis and M^-1 J^T from the bodies' FP32 inverse inertia
// (Bodies::iinv32), so a sweep reads half the bytes and a whole batch's
// records fit the 256 MB last-level cache.
#define CP_RS 64
#define CP_RSF 32
enum RowFieldF32 {
	CP_F_IMA = 11,  // 1/m_a (0 unless dynamic)
	CP_F_N = 16,    // normal [3]
	CP_F_IMB = 19,  // 1/m_b
	CP_F_U = 20,    // friction direction u [3]
	CP_F_W = 23,    // friction direction w [3]
	CP_F_RA = 26,   // contact point - centre of mass of body a [3]
	CP_F_RB = 29    // contact point - centre of mass of body b [3]
};
enum RowField {
	CP_R_IMP = 0,     // accumulated normal impulse
	CP_R_PIMP = 1,    // accumulated pseudo impulse
	CP_R_FIMP = 2,    // accumulated friction impulses [2]
	CP_R_SREADY = 4,  // friction cone not saturated in the last sweep (0/1)
	CP_R_MU = 5,      // friction coefficient
	CP_R_INV = 6,     // (J M^-1 J^T (1+cfm))^-1
	CP_R_TGT = 7,     // velocity target
	CP_R_PTGT = 8,    // pseudo-velocity target
	CP_R_FTGT = 9,    // friction targets [2]; 11 padding
	CP_R_FINV = 12,   // friction 2x2 inverse [4]
	CP_R_BASIS = 16,  // n[3], 1/m_a, u[3], 1/m_b, w[3], padding (1/m = 0 unless dynamic)
	CP_R_ANG = 28,    // normal row angular parts: J_a[3], J_b[3], (M_a^-1 J_a^T)[3], (M_b^-1 J_b^T)[3]
	CP_R_FANG = 40    // friction angular parts: J_a[2][3], J_b[2][3], (M_a^-1 J_a^T)[3][2], (M_b^-1 J_b^T)[3][2]
};

// A step's manifold points, compacted and sorted by pair key.
struct NewPts {
	uint64_t *key, *cid, *bid;
	double *pos, *nrm;  // [3][cap]
	double *pen, *mat;  // mat [3][cap]: restitution, mu_k, mu_s
	int *is_new, *src;
};

#define CP_MAXC 128  // colours per scene (plus one serial bucket)

// Coloured edge slots: one descriptor per position of a scene's colour-sorted
// edge list, so a sweep reads one 16-byte record per edge instead of a chain
// of dependent loads.  x = first cache index, y = contact count (low 16 bits)
// | ED_LONG (more than 32 contacts: live flags are read per contact) | dynamic
// flags of the two bodies, z / w = body indices; lmask bit j = contact x + j
// is live.
#define ED_LONG (1 << 29)
#define ED_DYN_A (1 << 30)
#define ED_DYN_B (1u << 31)

// Launchers (defined in the kernel translation units).
void launch_gravity(const Bodies& b, const Params& p, hipStream_t s);
void launch_refresh_geometry(const Bodies& b, const Shapes& sh, int only_moved, hipStream_t s);
void launch_update_inertia(const Bodies& b, hipStream_t s);
// Serving: body records (3 ints + BODY_REC doubles, see k_body_scatter)
// written into the SoA arrays with world inertia and geometry derived; list =
// the records' body indices.  Per-body scene index from scene offsets.
#define BODY_REC 35
#define SLOT_PTS 32  // world points reserved per body in a serving world
#define SLOT_NRM 64  // world Gauss normals reserved per body in a serving world
void launch_body_scatter(const Bodies& b, const Shapes& sh, int count, const int* ri, const double* rd, const int* list, hipStream_t s);
void launch_scene_of_body(int n, const int* sbo, int nscenes, int* out, hipStream_t s);
// Contacts whose first body lies in one of the ranges [r[2k], r[2k] + r[2k+1])
// are retired: not live, removed at the next merge.
void launch_kill_contacts(const Contacts& C, int n, const int* ranges, int nranges, hipStream_t s);
// The step-start loop with sleeping: movement history, sleep counters, gravity.
void launch_gravity_sleep(const Bodies& b, const Params& p, const SleepDev& S, hipStream_t s);
void launch_integrate(const Bodies& b, const Params& p, hipStream_t s);
// Flags dynamic bodies (of the visited prefix) with a non-finite pose or
// velocity: body_flag[i] = 1 and any[0] = 1.  Fill: p[0..n) = v.
void launch_check_finite(const Bodies& b, int* body_flag, int* any, hipStream_t s);
void launch_fill_int(int* p, int n, int v, hipStream_t s);
void launch_gather_states(const Bodies& b, const int* start, const int* off, int nsc, int rows, const int* body_flag, double* out, int* flag,
                          hipStream_t s);
// p[list[k]] = 0 for k < n.
void launch_zero_list(int* p, const int* list, int n, hipStream_t s);
void launch_bp_keys(const Bodies& b, uint64_t* keys, int* vals, hipStream_t s);
// skip_nonfinite: store non-finite boxes empty (they overlap nothing).
void launch_bp_gather(const Bodies& b, const int* sorted, double* sb, int skip_nonfinite, hipStream_t s);
#define BP_CH 64  // sweep positions per broadphase thread
void launch_bp_extent(int n, const int* sorted, const double* sb, const int* scene_of_body, const int* scene_body_off, int* ext,
                      int* nchunk, hipStream_t s);
void launch_bp_owner(int n, const int* choff, int* owner, hipStream_t s);
void launch_bp_sweep(int n, int nch, const int* sorted, const double* sb, const int* ext, const int* choff, const int* owner, int* counts,
                     const int* offs, uint64_t* out, int write, hipStream_t s);
// lists: 6 * npairs ints (per-class pair lists), counts: 8 ints.
// nocol: sorted pair keys (lo << 32 | hi) that never collide (jointed bodies);
// rank (nullable): slot rank of each pair (mesh pairs own MESH_SLOTS
// consecutive slots, rank 0 runs the test, the rest start empty).
void launch_narrowphase(const Bodies& b, const Shapes& sh, const Params& p, const uint64_t* pairs, int npairs, const Manifolds& m,
                        int cap, int* lists, int* counts, const uint64_t* nocol, int n_nocol, const int* rank, hipStream_t s, const Fork* fk = nullptr);
// Mesh pairs: slot expansion (count pass, then write pass at the scanned
// offsets), candidate triangles per pair (same two passes), one wave per
// candidate, then per pair the merge and cull into its slots.
void launch_mesh_expand(const Bodies& b, const MeshDev& md, const uint64_t* pairs, int npairs, const uint64_t* nocol, int n_nocol,
                        int* cnt, const int* off, uint64_t* out, int* rank, hipStream_t s);
void launch_mesh_candidates(const Bodies& b, const MeshDev& md, const uint64_t* pairs, const int* list, const int* counts, int nlist,
                            int* cnt, const int* off, MeshItems it, int* overflow, hipStream_t s);
void launch_mesh_triangles(const Bodies& b, const Shapes& sh, const MeshDev& md, const Params& p, const uint64_t* pairs, MeshItems it,
                           int nitems, hipStream_t s);
void launch_mesh_merge(const Params& p, const uint64_t* pairs, const MeshDev& md, const int* list, int nlist, const int* off, MeshItems it,
                       Manifolds m, int cap, int* overflow, hipStream_t s);
// Batched raycasts: per ray the first body hit among its scene's bodies
// (ray_scene[r] = -1 or a null ray_scene: every body); hit_body[r] (-1: none)
// and hit7[r] = point xyz, normal xyz, distance (written on a hit).
void launch_raycast(const Bodies& b, const Shapes& sh, const MeshDev& md, const int* scene_body_off, int n_rays, const double* org,
                    const double* dir, const int* ray_scene, int* hit_body, double* hit7, hipStream_t s);
// Class list index of mesh pairs in the narrowphase lists / counts.
#define NPC_MESH_LIST 6
void launch_np_gather(const Manifolds& M, int mcap, const uint64_t* pairs, int npairs, const int* off, const NewPts& np, int cap,
                      hipStream_t s);
void launch_np_match(const NewPts& np, int nnew, int cap, const Contacts& C, int nold, const Bodies& b, const Params& P,
                     hipStream_t s);
void launch_old_keep(const Contacts& C, int nold, int* keep, const int* prot, hipStream_t s);
void launch_new_keys(const NewPts& np, int nnew, const int* rank, uint64_t* nk, hipStream_t s);
void launch_merge(const Contacts& S, int nold, const int* keep_scan, const uint64_t* nk, int nnewc, const NewPts& np, int nnew, int cap,
                  const int* rank, const Contacts& D, const Bodies& b, const Params& P, uint64_t seq0, hipStream_t s);
void launch_age(const Contacts& C, int n, hipStream_t s, const int* ncp = nullptr);
void launch_set_count(int* p, int v, hipStream_t s);
void launch_edge_heads(const Contacts& C, int n, int* head, hipStream_t s);
void launch_edge_build(const Contacts& C, int n, const int* head, const int* eidx, uint64_t* ekey, int* efirst, int* ecnt, int* elive,
                       hipStream_t s);
void launch_edge_slots(const Bodies& b, const Contacts& C, const int* order, int nslots, const int* efirst, const int* ecnt, int4* desc,
                       uint32_t* lmask, hipStream_t s);
// Edge colouring of nscenes scenes (scene_list, or scenes 0..nscenes-1);
// max_bodies: the largest of them.
void launch_color(const Bodies& b, const int* scene_body_off, int nscenes, const uint64_t* ekey, const int* elive, int nedges,
                  int* ecolor, int* order, int* color_off, int* ncolors, int* edge_range, int max_bodies, hipStream_t s,
                  const int* scene_list = nullptr);
void launch_build_rows(const Bodies& b, const Contacts& C, int nc, const Params& P, hipStream_t s, const int* ncp = nullptr);
// The FP32 inverse inertia of every body (the split solve's rows read it).
void launch_iinv32(const Bodies& b, hipStream_t s);
// Joint rows: mode 0 all joints (holonomic ones skipped when first), 1 the
// holonomic joints with warm-start re-projection, 2 the holonomic joints
// without re-projection.
void launch_joint_init(const Bodies& b, const Joints& J, int k0, int k1, hipStream_t s);
// Springs at a substep's start: groups [goff[g], goff[g+1]) of list (joint
// indices), each group in the constraint graph's edge order, groups sharing no
// dynamic body.
void launch_springs(const Bodies& b, const Joints& J, int ngroups, const int* goff, const int* list, double sub_dt, hipStream_t s);
void launch_joint_rows(const Bodies& b, const Joints& J, const Params& P, double sub_dt, int mode, int first, hipStream_t s);
// Block LDL factor of every holonomic system.
void launch_hsys_factor(const Joints& J, const HSysDev& H, double cfm, hipStream_t s);
// rows: (index, kind) pairs, kind 0 contact normal, 1 friction, 2 joint;
// isl_sys/isl_sys_off: holonomic systems per island; lin: [6][n] scratch.
void launch_pgs_islands(const Bodies& b, const Contacts& C, const Joints& J, const HSysDev& H, const IslandSched& S, const Params& P,
                        int* applied, double* lin, int first, bool colored, hipStream_t s);
void launch_pgs_colored(const Bodies& b, const Contacts& C, const int* scene_body_off, int nscenes, const int* edge_range, const int4* desc,
                        const uint32_t* lmask, const int* color_off, const int* ncolors, const Params& P, int threads, hipStream_t s,
                        const int* scene_list, int nlist);
void launch_pgs_colored_lds(const Bodies& b, const Contacts& C, const int* scene_list, int nlist, int max_bodies, const int* scene_body_off,
                            const int* edge_range, const int4* desc, const uint32_t* lmask, const int* color_off, const int* ncolors,
                            const Params& P, int threads, hipStream_t s);
void launch_rows_writeback(const Contacts& C, int nc, const Params& P, hipStream_t s, const int* ncp = nullptr);
// The whole substep loop of each listed scene in one workgroup (rows, coloured
// solve, impulses back into the cache, integration; P.substeps times): the LDS
// variant for scenes of at most max_bodies bodies, the global-memory one for
// larger scenes.  efirst / ecnt: first cache index and contact count per edge.
void launch_substeps_lds(const Bodies& b, const Contacts& C, const int* scene_list, int nlist, int max_bodies, const int* scene_body_off,
                         const int* edge_range, const int* efirst, const int* ecnt, const int4* desc, const uint32_t* lmask,
                         const int* color_off, const int* ncolors, const Params& P, int threads, hipStream_t s);
void launch_substeps_global(const Bodies& b, const Contacts& C, const int* scene_list, int nlist, const int* scene_body_off, const int* edge_range,
                            const int* efirst, const int* ecnt, const int4* desc, const uint32_t* lmask, const int* color_off,
                            const int* ncolors, const Params& P, int threads, hipStream_t s);
void launch_pgs_residual(const Bodies& b, const Contacts& C, int nc, const Params& P, double* acc, hipStream_t s);
// rows: (index, kind) pairs, kind 0 contact normal, 1 friction, 2 joint;
// isl_sys/isl_sys_off: holonomic systems per island; lin: [6][n] scratch.
void launch_pgs_islands(const Bodies& b, const Contacts& C, const Joints& J, const HSysDev& H, const IslandSched& S, const Params& P,
                        int* applied, double* lin, int first, bool colored, hipStream_t s);
void launch_pgs_colored(const Bodies& b, const Contacts& C, const int* scene_body_off, int nscenes, const int* edge_range, const int* order,
                        const int* color_off, const int* ncolors, const int* efirst, const int* ecnt, const Params& P, int threads,
                        hipStream_t s, const int* scene_list = nullptr, int nlist = 0);

// Island search (k_islands.hip), in launch order.
void launch_ib_keys(const uint64_t* ekey, int ne, const uint64_t* jgkey, int ng, uint64_t* keys, int* vals, hipStream_t s);
void launch_ib_heads(const uint64_t* skeys, int n, int* head, hipStream_t s);
void launch_ib_edges(const uint64_t* skeys, const int* svals, int n, const int* head, const int* midx, int ne, const int* efirst,
                     const int* ecnt, const int* live, const int* jgoff, const int* jord, const JointAct& jact, int* me,
                     uint64_t* adjk, int* adjv, hipStream_t s);
void launch_ib_aoff(const uint64_t* adjk, int n2, int nb, int* aoff, hipStream_t s);
void launch_ib_components(const int* me, const int* midx, int n, int nb, const int* movement, const int* rank, const int* aoff, int* parent,
                          int* comp, IslCnt* acc, const SleepDev& S, hipStream_t s);
void launch_ib_flags(int nb, const int* rord, const int* comp, const IslCnt* acc, int sleeping, IslCnt* in, hipStream_t s);
void launch_ib_tables(int nb, const int* rord, const IslCnt* in, const IslCnt* out, const IslTables& T, int* info, const int* comp,
                      const IslCnt* acc, const SleepDev& S, double* vel, double* angv, hipStream_t s);
void launch_ib_walk(const IslWalk& W, int nisl, bool colored, hipStream_t s);

}  // namespace cpx
