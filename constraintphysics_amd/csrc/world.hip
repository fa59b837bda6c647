// Host runtime behind the C ABI (include/cp_abi.h): owns every device buffer
// of a world, uploads scenes, and issues the per-step launch sequence that
// replaces PhysicsEngine::timeStep (PhysicsEngine.cpp:50-471):
//
//   gravity -> broadphase (segmented radix sort by AABB min.x, sweep, pair sort)
//   -> narrowphase (wave per pair) -> point compaction -> contact persistence
//   (match + merge) -> constraint-graph edges -> schedule (edge colouring, or the
//   reference's island order built on the host) -> substeps x (row build, PGS,
//   integrate) -> contact ageing -> world-geometry refresh.
//
// Sizes that drive later launches (pair, point and contact counts) are read
// back with one small synchronous copy each; everything else stays on the
// device.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <limits>
#include <array>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/cp_abi.h"
#include "cp_internal.h"
#include "shapes.h"

using namespace cpx;

#define HIPCHK(x)                                  \
	do {                                           \
		hipError_t e_ = (x);                       \
		if (e_ != hipSuccess) throw DevErr{e_, #x}; \
	} while (0)

namespace {

struct DevErr {
	hipError_t e;
	const char* what;
};
struct ApiErr {
	int code;
};

// Grow-only device array.
template <class T>
struct DArr {
	T* p = nullptr;
	size_t cap = 0;
	void reserve(size_t n) {
		if (n <= cap && p) return;
		size_t m = std::max<size_t>(n, 16);
		if (p) HIPCHK(hipFree(p));
		p = nullptr;
		if (hipMalloc(&p, m * sizeof(T)) != hipSuccess) {
			p = nullptr;
			cap = 0;
			throw ApiErr{CP_ERR_OOM};
		}
		cap = m;
	}
	void free_() {
		if (p) (void)hipFree(p);
		p = nullptr;
		cap = 0;
	}
};

struct ContactBuf {
	int cap = 0;
	std::vector<void*> allocs;
	Contacts c{};
	void release() {
		for (void* a : allocs) (void)hipFree(a);
		allocs.clear();
		cap = 0;
		c = Contacts{};
	}
	template <class T>
	T* al(size_t n) {
		T* p = nullptr;
		if (hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) throw ApiErr{CP_ERR_OOM};
		allocs.push_back(p);
		return p;
	}
	void reserve(int n) {
		if (n <= cap) return;
		release();
		int K = std::max(n + n / 2, 256);
		cap = K;
		c.cap = K;
		c.key = al<uint64_t>(K); c.cid = al<uint64_t>(K); c.bid = al<uint64_t>(K); c.seq = al<uint64_t>(K);
		c.lp1 = al<double>(3 * K); c.lp2 = al<double>(3 * K); c.ln1 = al<double>(3 * K);
		c.pen = al<double>(K); c.bounce = al<double>(K); c.mu = al<double>(K);
		c.life = al<int>(K); c.live = al<int>(K); c.sready = al<int>(K);
		c.imp_n = al<double>(K); c.imp_t = al<double>(2 * K);
		c.fu = al<double>(3 * K); c.fw = al<double>(3 * K);
		c.row = al<double>((size_t)CP_RS * K);
	}
};

template <class T>
void up(DArr<T>& d, const std::vector<T>& h) {
	d.reserve(h.size());
	if (!h.empty()) HIPCHK(hipMemcpy(d.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
}
template <class T>
void down(std::vector<T>& h, const T* d, size_t n) {
	h.resize(n);
	if (n) HIPCHK(hipMemcpy(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost));
}
template <class T>
T read1(const T* d) {
	T v;
	HIPCHK(hipMemcpy(&v, d, sizeof(T), hipMemcpyDeviceToHost));
	return v;
}

// Grows d to hold n elements keeping its first `keep` elements.
template <class T>
void grow_keep(DArr<T>& d, size_t n, size_t keep) {
	if (n <= d.cap && d.p) return;
	size_t m = std::max<size_t>(n + n / 2, 16);
	T* p = nullptr;
	if (hipMalloc(&p, m * sizeof(T)) != hipSuccess) throw ApiErr{CP_ERR_OOM};
	if (keep && d.p) HIPCHK(hipMemcpy(p, d.p, keep * sizeof(T), hipMemcpyDeviceToDevice));
	if (d.p) HIPCHK(hipFree(d.p));
	d.p = p;
	d.cap = m;
}
// Appends h[done..] to the device copy (host tables only ever grow).
template <class T>
void up_tail(DArr<T>& d, const std::vector<T>& h, size_t& done) {
	grow_keep(d, h.size(), done);
	if (h.size() > done) HIPCHK(hipMemcpy(d.p + done, h.data() + done, (h.size() - done) * sizeof(T), hipMemcpyHostToDevice));
	done = h.size();
}

}  // namespace

// Largest island (bodies + the zero column) whose velocity deltas the
// coloured-island solve keeps in LDS: 12 doubles per column, 48 KB.
#define ISL_LDS_BODIES 512

struct cp_world {
	cp_config cfg;
	hipStream_t stream = nullptr;
	// the PGS size classes run concurrently: two side streams forked from and
	// joined back into `stream` around each solve (default priority)
	hipStream_t side[2] = {nullptr, nullptr};
	hipEvent_t fork_ev = nullptr, join_ev[2] = {nullptr, nullptr};
	Fork fk{};
	// captured substep loop of the coloured mode, one per contact buffer
	struct Captured {
		hipGraphExec_t exec = nullptr;
		std::vector<uint64_t> key;
	};
	Captured graphs[2];
	DArr<int> d_nc;  // the step's contact count on the device (read by the captured kernels)
	int device = 0;
	// host copies of static data
	HostShapes hs;
	std::vector<int> b_shape, b_move, b_scene;
	std::vector<double> b_massprops, b_mat;
	std::vector<int64_t> node_order;
	// device: shapes
	DArr<int> s_type, s_pt_off, s_face_off, s_arc_off, s_edge_off, s_fv_off, s_fv, s_ncalc, s_ndir, s_sid, s_red, s_minpid, s_maxpid,
	    s_refidx, s_arcs, s_edges, s_vaf_off, s_vaf, s_vae_off, s_vae;
	DArr<double> s_ref_pts, s_gnorm, s_gmin, s_gmax, s_grefval, s_radius, s_height;
	// device: bodies
	DArr<double> com, q, vel, angv, pvel, pang, prev_com, prev_q, I, Iinv, Iref, Irefinv, mass, mat, aabb, wpts, wnorm, dv;
	DArr<float> iinv32;
	DArr<int4> bgeo;
	DArr<int> movement, shape, scene, wpt_off, wgv_off, color_mask, scene_body_off, scene_of_body;
	int nb = 0, nscenes = 0;
	bool uploaded = false;
	// step temporaries
	DArr<uint64_t> bp_keys, bp_keys2, pairs, pairs2, nk;
	DArr<int> bp_vals, bp_vals2, bp_counts, bp_ext, bp_choff, bp_owner, bp_ccount, bp_coffs;
	DArr<double> bp_sb;
	DArr<char> tmp;
	int mcap = 0;
	Manifolds man{};
	std::vector<void*> man_allocs;
	DArr<int> pt_cnt, pt_off, np_lists, np_counts;
	int npcap = 0;
	NewPts np{};
	std::vector<void*> np_allocs;
	DArr<int> keep, keep_scan, rank;
	ContactBuf cb[2];
	int cur = 0;
	int ncontacts = 0;
	uint64_t seq = 1ull << 32;
	DArr<int> e_head, e_idx, e_first, e_cnt, e_live, e_color, e_order, color_off, ncolors, edge_range, applied, isl_rows, isl_off;
	DArr<uint64_t> e_key;
	DArr<int4> e_desc;
	DArr<uint32_t> e_lmask;
	int nedges = 0;
	int live_contacts = 0;
	DArr<double> resid_acc;
	// non-finite detection: per-body flags (sticky until the state is set),
	// the step's any-flag and its pinned host copy
	DArr<int> nf_body, nf_any;
	int* nf_host = nullptr;
	bool nf_step = false;  // the last cp_world_step call saw a non-finite body
	std::vector<std::vector<int>> cls_lists;
	std::vector<std::pair<int, int>> cls_sizes;  // (scene, bodies) the classes are built from
	int cls_built[3] = {0, 0, 0};                 // class bounds the lists were built with
	int max_scene_bodies = 0;                     // largest scene the classes hold
	void build_classes();
	int cls_max[4] = {0, 0, 0, 0};
	DArr<int> cls_dev[4];
	double last_resid[4] = {0, 0, 0, 0};
	std::vector<int64_t> last_rows, last_offs, last_scene_coff;
	hipEvent_t ev[10];
	hipEvent_t pgs_ev[2 * 64];
	double last_ms[8] = {0};
	bool timing = true;
	bool fused_step = false;  // the last step ran the fused substep kernels
	// joints (host copies of static data; device SoA)
	std::vector<int> j_type, j_kind, j_a, j_b;  // j_type: degree (rows); j_kind: cp_joint_add's type
	std::vector<double> j_geo;  // 14 per joint: tp1, tp2, ax1, ax2, ps, rs
	std::vector<double> j_drive;  // 4 per joint: powered mode, max, target velocity, target position (motors, piston drives)
	std::vector<double> j_cfm;    // per joint: custom CFM, NaN: the world's
	// springs: groups of springs sharing dynamic bodies, each in the
	// constraint graph's edge order, for the joint count and root order they
	// were built for
	DArr<int> sp_goff, sp_list;
	int sp_ngroups = 0, sp_nj = -1, n_springs = 0;
	std::vector<int> sp_roots;
	bool drive_dirty = false;
	int j_dev_n = 0;            // joints on the device
	bool joints_dirty = false;  // static joint data or systems changed since the upload
	DArr<int> dj_type, dj_kind, dj_a, dj_b, dj_sys, dj_proj;
	DArr<double> dj_tp1, dj_tp2, dj_ax1, dj_ax2, dj_ps, dj_rs, dj_imp, dj_pimp, dj_himp, dj_u, dj_w, dj_row, dj_drive, dj_cfm, lin;
	DArr<uint64_t> nocol;
	int n_nocol = 0;
	// Joint edges (body pairs with joints) of the constraint graph: n1 is b1 of
	// the pair's first joint (getOrCreateEdgeTo, PhysicsEngine.cpp:1958-1971).
	struct JEdge {
		int n1, n2;
		std::vector<int> joints;  // creation order
		int nhol = 0;             // holonomic joints among them (hasHolonomicConstraint)
		bool scan = true;         // holonomic_system_scan_needed
		int sys = -1;
	};
	std::map<uint64_t, JEdge> jedges;
	std::vector<std::vector<uint64_t>> sys_edges;  // member edges per system
	bool hs_explicit = false;  // systems set by the caller: no discovery
	bool hs_changed = false;   // a joint joined an edge already in a system
	bool hs_scan = false;      // joints added since the last search
	// holonomic systems: constraint order (og), then the symbolic layout
	std::vector<std::vector<int>> hs_og;
	struct HLayout {
		std::vector<int> order, table, vloc;
		int degree = 0, cap = 0;
	};
	std::vector<HLayout> hs_lay;
	DArr<int> dh_nc, dh_ord_off, dh_order, dh_vloc, dh_tab_off, dh_table, dh_degree, dh_buf_off, dh_dbuf_off, isl_sys;
	DArr<int> dh_cp_off, dh_cs_ptr, dh_cs_row, dh_cs_loc, dh_dg_loc;
	DArr<double> dh_buf, dh_dbuf, dh_scratch;
	bool hs_pending = true;
	int hs_lds_doubles = 0;  // largest system's buffer + column scratch, in doubles
	// island schedule (built on the device each step of the island modes)
	DArr<int> isl_jlin, isl_eoff, isl_coff, isl_icoff, isl_ncol, isl_serial, isl_bod, isl_nbod, isl_cloc, isl_jloc, isl_sys_cnt;
	DArr<int> isl_root, isl_edge_off, isl_body_off, isl_body_cnt, isl_stack_off, isl_joint_off;
	DArr<uint64_t> ib_keys, ib_keys2, ib_adjk, ib_adjk2, ib_jgkey, ib_col_used;
	DArr<int> ib_vals, ib_vals2, ib_head, ib_midx, ib_me, ib_adjv, ib_adjv2, ib_aoff, ib_parent, ib_comp, ib_jgoff, ib_jord, ib_rank, ib_rord,
	    ib_info, ib_stack, ib_vlist, ib_elist, ib_vis, ib_ecol, ib_eorder, ib_loc, ib_sys_seen, ib_slots;
	DArr<IslCnt> ib_acc, ib_in, ib_out;
	int ib_nj = -1, ib_ng = 0;        // joints the pair groups were built from, groups
	std::vector<int> ib_rord_host;     // node-map order the ranks were uploaded from
	// sleeping: per-body state (allocated with the bodies), the movement
	// history ring, pending waking actions
	DArr<int> sl_asleep, sl_asleep_rm, sl_ticks, sl_hlen, sl_hhead, sl_disabled, sl_proc;
	DArr<double> sl_counter, sl_hist;
	int sl_hcap = -1;                  // history slots the ring holds
	bool sl_active = false;            // sleeping has been on: asleep flags are honoured
	std::vector<int> sl_wake;          // bodies with a waking action since the last step
	std::vector<int> sl_dis_host;      // sleep_disabled per body
	void prepare_sleep(bool& proc);
	// per cached contact: removed by the reference at the last step's first
	// substep (memory_life below -1 now, neither body asleep then)
	std::vector<char> removed_contacts(const std::vector<int>& life) {
		std::vector<char> out(life.size(), 0);
		std::vector<int> rm;
		std::vector<uint64_t> key;
		if (sl_active) {
			down(rm, sl_asleep_rm.p, nb);
			down(key, cb[cur].c.key, life.size());
		}
		for (size_t k = 0; k < life.size(); k++) {
			if (life[k] >= -1) continue;
			out[k] = !sl_active || (!rm[(size_t)(key[k] >> 32)] && !rm[(size_t)(key[k] & 0xFFFFFFFFu)]);
		}
		return out;
	}
	int isl_n = 0;                     // islands of the last schedule
	bool islands_on_device = false;    // last_rows / last_offs not yet read back
	int isl_lds_bodies = 0;            // LDS columns per island solve (largest island + 1, capped)
	int lin_nl = 0;
	IslTables isl_tables();
	void fetch_islands();
	SleepDev sleep_dev();
	// static meshes: one world instance per mesh body (triangles and AABB tree
	// placed on the host from the body's pose), device copies, step scratch
	std::vector<int> mesh_inst_of_body, mesh_body;
	bool meshes_placed = false;
	DArr<int> dm_inst_of_body, dm_root, dm_tri_i, dm_lr, pair_rank, mesh_cnt, mesh_off, mesh_ccnt, mesh_coff, mesh_overflow;
	DArr<double> dm_tri, dm_box;
	DArr<int> mi_slot, mi_tri, mi_count, mi_owner;
	DArr<double> mi_head, mi_pts, mi_mat;
	DArr<uint64_t> mi_magic, mi_avail;
	bool last_expanded = false;
	int mesh_depth = 0;
	// serving: a fixed body capacity; scenes are admitted into free ranges and
	// retired back to parked bodies without moving other scenes' bodies, so a
	// scene steps exactly as it would alone.  Changes are staged and applied
	// at the next step (or commit) in one scatter.
	bool serving = false;
	int park_shape = -1;
	struct Range {
		int start, len, handle;  // handle -1: free (parked bodies)
	};
	std::vector<Range> ranges;
	int next_handle = 0;
	DArr<int> gs_idx, gs_flag;  // cp_world_get_scene_states: scene starts and row offsets, non-finite flags
	DArr<double> gs_rows;       // its gathered rows
	std::vector<int> stage_i, kill_ranges, stage_mark;
	std::vector<double> stage_d;
	int commit_id = 0;
	bool serving_dirty = false;
	size_t sh_done[32] = {0};
	DArr<int> st_i, st_list, st_kill;
	DArr<double> ray_io;  // raycast inputs (origins, directions) and outputs
	DArr<int> ray_int;    // raycast scenes and hit bodies
	DArr<double> st_d;
	// admissions that resume a scene snapshot: world inertia and contact
	// cache rows (global ids) applied at the commit that scatters the bodies
	struct Resume {
		int start, n;
		std::vector<double> inertia;  // [n][18]
		std::vector<int64_t> ci;      // [nc][6]
		std::vector<uint64_t> cm;     // [nc][2]
		std::vector<double> cf;       // [nc][21]
	};
	std::vector<Resume> resumes;
	void apply_resumes();
	// trajectory recording: one frame of 7 doubles (com xyz, q r i j k) per
	// body after every step, the layout of the reference's recorded runs
	FILE* traj = nullptr;
	double* traj_host = nullptr;  // pinned staging (com [3][n], q [4][n])
	size_t traj_cap = 0;
	std::vector<double> traj_frame;
	void record_frame();
	void sync_shapes();
	void stage_body(int i, int shape, int move, const double* mp19, const double* mat4, const double* st13);
	void park_range(int first, int count);
	void serving_commit();

	int live_n = 0, live_scenes = 0;  // serving: occupied prefix (bodies, scenes)
	Bodies bodies() {
		Bodies b{};
		b.n = nb;
		b.nl = serving ? live_n : nb;
		b.com = com.p; b.q = q.p; b.vel = vel.p; b.angv = angv.p; b.pvel = pvel.p; b.pang = pang.p;
		b.prev_com = prev_com.p; b.prev_q = prev_q.p;
		b.I = I.p; b.Iinv = Iinv.p; b.Iref = Iref.p; b.Irefinv = Irefinv.p;
		b.mass = mass.p;
		b.movement = movement.p; b.shape = shape.p; b.scene = scene.p;
		b.mat = mat.p; b.aabb = aabb.p;
		b.wpt_off = wpt_off.p; b.wpts = wpts.p; b.wgv_off = wgv_off.p; b.wnorm = wnorm.p;
		b.dv = dv.p; b.iinv32 = iinv32.p; b.geo = bgeo.p; b.color_mask = color_mask.p;
		b.asleep = sl_active ? sl_asleep.p : nullptr;
		b.asleep_rm = sl_active ? sl_asleep_rm.p : nullptr;
		return b;
	}
	Shapes shapes() {
		Shapes s{};
		s.n = (int)hs.type.size();
		s.type = s_type.p; s.pt_off = s_pt_off.p; s.face_off = s_face_off.p; s.arc_off = s_arc_off.p; s.edge_off = s_edge_off.p;
		s.ref_pts = s_ref_pts.p; s.fv_off = s_fv_off.p; s.fv = s_fv.p;
		s.f_ncalc = s_ncalc.p; s.f_ndir = s_ndir.p; s.f_sid = s_sid.p; s.g_norm = s_gnorm.p;
		s.g_red = s_red.p; s.g_minpid = s_minpid.p; s.g_maxpid = s_maxpid.p; s.g_refidx = s_refidx.p;
		s.g_min = s_gmin.p; s.g_max = s_gmax.p; s.g_refval = s_grefval.p;
		s.arcs = s_arcs.p; s.edges = s_edges.p;
		s.va_face_off = s_vaf_off.p; s.va_faces = s_vaf.p; s.va_edge_off = s_vae_off.p; s.va_edges = s_vae.p;
		s.radius = s_radius.p; s.height = s_height.p;
		return s;
	}
	Params params() {
		Params p{};
		p.dt = cfg.dt;
		p.sub_dt = cfg.dt / cfg.substeps;
		p.gx = cfg.gravity[0]; p.gy = cfg.gravity[1]; p.gz = cfg.gravity[2];
		p.cfm = cfg.cfm;
		p.hardness = cfg.contact_hardness;
		double gm = std::sqrt(p.gx * p.gx + p.gy * p.gy + p.gz * p.gz);
		p.cutoff_vel = 2 * gm * cfg.dt;  // PhysicsEngine.h:193
		const double ref_pi = 3.14159265358979323846;  // CollisionDetect.h:20
		double tol = 1 * ref_pi / 180;
		p.cos_tol = 1 - std::cos(tol);
		p.sin_tol = std::sin(tol);
		for (int i = 0; i < 4; i++) {
			p.cull_c[i] = std::cos(2 * ref_pi * i / 4);
			p.cull_s[i] = std::sin(2 * ref_pi * i / 4);
		}
		p.vel_iters = cfg.vel_iters;
		p.pos_iters = cfg.pos_iters;
		p.substeps = cfg.substeps;
		p.contact_life = cfg.contact_life;
		p.gyro_ticks = cfg.gyro_ticks;
		p.defect_fixes = cfg.defect_fixes;
		p.precision = cfg.pgs_mode == CP_PGS_REFERENCE_ORDER ? CP_PREC_FP64 : cfg.precision;  // the parity mode stays FP64
		p.hol_iters = cfg.hol_iters;
		return p;
	}

	Joints joints() {
		Joints J{};
		J.n = (int)j_type.size();
		J.type = dj_type.p; J.kind = dj_kind.p; J.a = dj_a.p; J.b = dj_b.p; J.sys = dj_sys.p;
		J.tp1 = dj_tp1.p; J.tp2 = dj_tp2.p; J.ax1 = dj_ax1.p; J.ax2 = dj_ax2.p; J.ps = dj_ps.p; J.rs = dj_rs.p;
		J.imp = dj_imp.p; J.pimp = dj_pimp.p; J.himp = dj_himp.p; J.u = dj_u.p; J.w = dj_w.p; J.row = dj_row.p;
		J.proj = dj_proj.p; J.drive = dj_drive.p; J.cfm = dj_cfm.p;
		return J;
	}
	JointAct joint_act() { return JointAct{(int)j_type.size(), dj_kind.p, dj_u.p, dj_drive.p, dj_tp1.p}; }
	HSysDev hsys() {
		HSysDev H{};
		H.nsys = (int)hs_lay.size();
		H.nc = dh_nc.p; H.ord_off = dh_ord_off.p; H.order = dh_order.p; H.vloc = dh_vloc.p; H.tab_off = dh_tab_off.p; H.table = dh_table.p;
		H.degree = dh_degree.p; H.buf_off = dh_buf_off.p; H.dbuf_off = dh_dbuf_off.p; H.buf = dh_buf.p; H.dbuf = dh_dbuf.p;
		H.scratch = dh_scratch.p;
		H.cp_off = dh_cp_off.p; H.cs_ptr = dh_cs_ptr.p; H.cs_row = dh_cs_row.p; H.cs_loc = dh_cs_loc.p; H.dg_loc = dh_dg_loc.p;
		H.lds_doubles = hs_lds_doubles;
		return H;
	}
	MeshDev meshdev() {
		MeshDev m{};
		m.n_inst = (int)mesh_body.size();
		m.inst_of_body = dm_inst_of_body.p;
		m.root = dm_root.p;
		m.tri = dm_tri.p;
		m.tri_i = dm_tri_i.p;
		m.node_box = dm_box.p;
		m.node_lr = dm_lr.p;
		return m;
	}
	MeshItems meshitems() {
		return MeshItems{mi_slot.p, mi_tri.p, mi_count.p, mi_head.p, mi_pts.p, mi_magic.p, mi_mat.p, mi_owner.p, mi_avail.p};
	}
	void place_meshes(const double* state13);
	int mesh_narrowphase(int npairs);
	void upload_joints();
	void upload_drives();
	void init_new_joints(const Bodies& b);
	void schedule_springs();
	static int state_projection(int kind, const double* s6, const double* g);
	int j_init_from = 0;  // joints from this index on take their constructor state from the bodies
	void read_joint_state(double* s, int n);
	void write_joint_state(const double* s);
	void layout_systems();
	void discover_systems();
	const std::vector<int>& root_order();  // the reference's node-map order (cached per body count)
	std::vector<int> root_cache;
	int root_cache_nb = -1;
	void free_all();
	void upload();
	void ensure_manifolds(int n);
	void ensure_points(int n);
	void scan(const int* in, int* out, int n);
	int broadphase();
	int narrowphase(int npairs);
	void persist(int npairs);
	void edges();
	void schedule_islands(bool colored);
	void record_colored_schedule();
	void launch_pgs_colored_lds_global(const Bodies& b, const Params& P);
	void color_by_class(const Bodies& b);
	void step();
	void colored_pgs(const Bodies& b, const Params& P);
	void colored_substep(const Bodies& b, const Params& P, int s, const int* ncp);
	void fused_substeps(const Bodies& b, const Params& P);
	void graph_substeps(const Bodies& b, const Params& P);
};

void cp_world::free_all() {
	DArr<int>* ints[] = {&s_type, &s_pt_off, &s_face_off, &s_arc_off, &s_edge_off, &s_fv_off, &s_fv, &s_ncalc, &s_ndir, &s_sid, &s_red,
	                     &s_minpid, &s_maxpid, &s_refidx, &s_arcs, &s_edges, &s_vaf_off, &s_vaf, &s_vae_off, &s_vae, &movement, &shape,
	                     &scene, &wpt_off, &wgv_off, &color_mask, &scene_body_off, &scene_of_body, &bp_vals, &bp_vals2, &bp_counts,
	                     &bp_ext, &bp_choff, &bp_owner, &bp_ccount, &bp_coffs, &pt_cnt, &pt_off, &np_lists, &np_counts, &keep, &keep_scan, &rank, &e_head, &e_idx, &e_first, &e_cnt, &e_live, &e_color,
	                     &e_order, &color_off, &ncolors, &edge_range, &applied, &isl_rows, &isl_off};
	for (auto* a : ints) a->free_();
	for (auto& a : cls_dev) a.free_();
	DArr<double>* dbl[] = {&s_ref_pts, &s_gnorm, &s_gmin, &s_gmax, &s_grefval, &s_radius, &s_height, &com, &q, &vel, &angv, &pvel, &pang,
	                       &prev_com, &prev_q, &I, &Iinv, &Iref, &Irefinv, &mass, &mat, &aabb, &wpts, &wnorm, &dv, &bp_sb};
	for (auto* a : dbl) a->free_();
	iinv32.free_();
	bgeo.free_();
	gs_idx.free_();
	gs_flag.free_();
	gs_rows.free_();
	resid_acc.free_();
	ray_io.free_();
	ray_int.free_();
	DArr<uint64_t>* u64s[] = {&bp_keys, &bp_keys2, &pairs, &pairs2, &nk, &e_key};
	for (auto* a : u64s) a->free_();
	e_desc.free_();
	e_lmask.free_();
	d_nc.free_();
	nf_body.free_();
	nf_any.free_();
	DArr<int>* jints[] = {&dj_type, &dj_kind, &dj_a, &dj_b, &dj_sys, &dj_proj, &dh_nc, &dh_ord_off, &dh_order, &dh_vloc, &dh_tab_off, &dh_table, &dh_degree,
	                      &dh_buf_off, &dh_dbuf_off, &isl_sys, &isl_jlin, &isl_eoff, &isl_coff, &isl_icoff, &isl_ncol,
	                      &isl_serial, &isl_bod, &isl_nbod, &isl_cloc, &isl_jloc, &isl_sys_cnt, &dh_cp_off, &dh_cs_ptr,
	                      &dh_cs_row, &dh_cs_loc, &dh_dg_loc, &isl_root, &isl_edge_off, &isl_body_off, &isl_body_cnt, &isl_stack_off,
	                      &isl_joint_off, &ib_vals, &ib_vals2, &ib_head, &ib_midx, &ib_me, &ib_adjv, &ib_adjv2, &ib_aoff, &ib_parent,
	                      &ib_comp, &ib_jgoff, &ib_jord, &ib_rank, &ib_rord, &ib_info, &ib_stack, &ib_vlist, &ib_elist, &ib_vis,
	                      &ib_ecol, &ib_eorder, &ib_loc, &ib_sys_seen, &ib_slots, &sp_goff, &sp_list};
	for (auto* a : jints) a->free_();
	DArr<uint64_t>* ju64[] = {&ib_keys, &ib_keys2, &ib_adjk, &ib_adjk2, &ib_jgkey, &ib_col_used};
	for (auto* a : ju64) a->free_();
	ib_acc.free_();
	ib_in.free_();
	ib_out.free_();
	ib_nj = -1;
	ib_rord_host.clear();
	sp_nj = -1;
	sp_ngroups = 0;
	sp_roots.clear();
	isl_n = 0;
	islands_on_device = false;
	DArr<double>* jdbl[] = {&dj_tp1, &dj_tp2, &dj_ax1, &dj_ax2, &dj_ps, &dj_rs, &dj_imp, &dj_pimp, &dj_himp, &dj_u, &dj_w, &dj_row, &dj_drive, &dj_cfm, &lin,
	                        &dh_buf, &dh_dbuf, &dh_scratch};
	for (auto* a : jdbl) a->free_();
	nocol.free_();
	j_dev_n = 0;
	DArr<int>* mints[] = {&dm_inst_of_body, &dm_root, &dm_tri_i, &dm_lr, &pair_rank, &mesh_cnt, &mesh_off, &mesh_ccnt, &mesh_coff,
	                      &mesh_overflow, &mi_slot, &mi_tri, &mi_count, &mi_owner};
	for (auto* a : mints) a->free_();
	DArr<double>* mdbl[] = {&dm_tri, &dm_box, &mi_head, &mi_pts, &mi_mat};
	for (auto* a : mdbl) a->free_();
	mi_magic.free_();
	mi_avail.free_();
	tmp.free_();
	for (void* a : man_allocs) (void)hipFree(a);
	man_allocs.clear();
	for (void* a : np_allocs) (void)hipFree(a);
	np_allocs.clear();
	cb[0].release();
	cb[1].release();
}

// Upload static scene data (shapes, bodies) and size per-body buffers.
void cp_world::upload() {
	const HostShapes& h = hs;
	up(s_type, h.type); up(s_pt_off, h.pt_off); up(s_face_off, h.face_off); up(s_arc_off, h.arc_off); up(s_edge_off, h.edge_off);
	up(s_ref_pts, h.ref_pts); up(s_fv_off, h.fv_off); up(s_fv, h.fv); up(s_ncalc, h.ncalc); up(s_ndir, h.ndir); up(s_sid, h.sid);
	up(s_gnorm, h.gnorm); up(s_red, h.red); up(s_minpid, h.minpid); up(s_maxpid, h.maxpid); up(s_refidx, h.refidx);
	up(s_gmin, h.gmin); up(s_gmax, h.gmax); up(s_grefval, h.grefval); up(s_arcs, h.arcs); up(s_edges, h.edges);
	up(s_vaf_off, h.vaf_off); up(s_vaf, h.vaf); up(s_vae_off, h.vae_off); up(s_vae, h.vae);
	up(s_radius, h.radius); up(s_height, h.height);
	int n = nb;
	std::vector<double> Ir(9 * (size_t)n), Iri(9 * (size_t)n), ms(n), mt(3 * (size_t)n);
	std::vector<int> wpo(n + 1), wgo(n + 1);
	for (int i = 0; i < n; i++) {
		const double* mp = &b_massprops[19 * (size_t)i];
		ms[i] = mp[0];
		for (int c = 0; c < 9; c++) {
			Ir[c * (size_t)n + i] = mp[1 + c];
			Iri[c * (size_t)n + i] = mp[10 + c];
		}
		for (int c = 0; c < 3; c++) mt[c * (size_t)n + i] = b_mat[4 * (size_t)i + 1 + c];
		int s = b_shape[i];
		// world points and normals per body: hull vertices and Gauss normals;
		// sphere centre; cylinder centre + cap polygons and axis + Gauss vertices
		wpo[i + 1] = wpo[i] + (serving ? SLOT_PTS : h.pt_off[s + 1] - h.pt_off[s]);
		wgo[i + 1] = wgo[i] + (serving ? SLOT_NRM : h.face_off[s + 1] - h.face_off[s]);
	}
	up(Iref, Ir); up(Irefinv, Iri); up(mass, ms); up(mat, mt);
	up(movement, b_move); up(shape, b_shape); up(scene, b_scene);
	up(wpt_off, wpo); up(wgv_off, wgo);
	wpts.reserve(3 * (size_t)wpo[n]);
	wnorm.reserve(3 * (size_t)wgo[n]);
	// scenes
	std::vector<int> sbo;
	nscenes = 0;
	for (int i = 0; i < n; i++) {
		if (i == 0 || b_scene[i] != b_scene[i - 1]) {
			sbo.push_back(i);
			nscenes++;
		}
	}
	sbo.push_back(n);
	std::vector<int> sob(n);
	for (int s = 0; s < nscenes; s++)
		for (int i = sbo[s]; i < sbo[s + 1]; i++) sob[i] = s;
	up(scene_body_off, sbo);
	up(scene_of_body, sob);
	cls_sizes.clear();
	for (int s = 0; s < nscenes; s++) cls_sizes.push_back({s, sbo[s + 1] - sbo[s]});
	build_classes();
	DArr<double>* per3[] = {&com, &vel, &angv, &pvel, &pang, &prev_com};
	for (auto* a : per3) {
		a->reserve(3 * (size_t)n);
		HIPCHK(hipMemset(a->p, 0, 3 * (size_t)n * sizeof(double)));
	}
	q.reserve(4 * (size_t)n);
	prev_q.reserve(4 * (size_t)n);
	I.reserve(9 * (size_t)n);
	Iinv.reserve(9 * (size_t)n);
	aabb.reserve(6 * (size_t)n);
	dv.reserve(12 * (size_t)n);
	iinv32.reserve(8 * (size_t)n);
	bgeo.reserve(3 * (size_t)n);
	color_mask.reserve(4 * (size_t)n);
	applied.reserve(n);
	DArr<int>* sl_int[] = {&sl_asleep, &sl_asleep_rm, &sl_ticks, &sl_hlen, &sl_hhead, &sl_proc};
	for (auto* a : sl_int) {
		a->reserve(n);
		HIPCHK(hipMemset(a->p, 0, (size_t)n * sizeof(int)));
	}
	sl_counter.reserve(n);
	HIPCHK(hipMemset(sl_counter.p, 0, (size_t)n * sizeof(double)));
	sl_dis_host.resize(n, 0);
	up(sl_disabled, sl_dis_host);
	sl_hcap = -1;
	nf_body.reserve(n);
	HIPCHK(hipMemset(nf_body.p, 0, (size_t)n * sizeof(int)));
	nf_any.reserve(1);
	// mesh bodies (placed when their pose is set)
	mesh_inst_of_body.assign(n, -1);
	mesh_body.clear();
	for (int i = 0; i < n; i++)
		if (hs.type[b_shape[i]] == SHAPE_MESH) {
			mesh_inst_of_body[i] = (int)mesh_body.size();
			mesh_body.push_back(i);
		}
	up(dm_inst_of_body, mesh_inst_of_body);
	meshes_placed = mesh_body.empty();
	uploaded = true;
}

// ---------------------------------------------------------------- joints
// Joint state per joint: CP_JMAX accumulated impulses (unused components
// 0), then the basis u, w of the current rows (hinges, sliders, sliding
// hinges); a limit keeps is_inactive and the twist's tracked angle in u.
void cp_world::read_joint_state(double* s, int n) {
	int K = j_dev_n;
	std::vector<double> imp, u, w;
	down(imp, dj_imp.p, CP_JMAX * (size_t)K);
	down(u, dj_u.p, 3 * (size_t)K);
	down(w, dj_w.p, 3 * (size_t)K);
	for (int k = 0; k < n && k < K; k++) {
		double* r = s + CP_JSTATE * (size_t)k;
		for (int c = 0; c < CP_JMAX; c++) r[c] = c < j_type[k] ? imp[c * (size_t)K + k] : 0.0;
		for (int c = 0; c < 3; c++) {
			r[CP_JMAX + c] = u[c * (size_t)K + k];
			r[CP_JMAX + 3 + c] = w[c * (size_t)K + k];
		}
	}
}
void cp_world::write_joint_state(const double* s) {
	int K = j_dev_n;
	std::vector<double> imp(CP_JMAX * (size_t)K), u(3 * (size_t)K), w(3 * (size_t)K), zero(CP_JMAX * (size_t)K, 0.0);
	std::vector<int> proj(K, 0);
	for (int k = 0; k < K; k++) {
		const double* r = s + CP_JSTATE * (size_t)k;
		for (int c = 0; c < CP_JMAX; c++) imp[c * (size_t)K + k] = r[c];
		for (int c = 0; c < 3; c++) {
			u[c * (size_t)K + k] = r[CP_JMAX + c];
			w[c * (size_t)K + k] = r[CP_JMAX + 3 + c];
		}
		if (joint_is_inequality(j_kind[k])) proj[k] = state_projection(j_kind[k], r + CP_JMAX, &j_geo[14 * (size_t)k]);
	}
	up(dj_imp, imp); up(dj_u, u); up(dj_w, w); up(dj_pimp, zero); up(dj_himp, zero); up(dj_proj, proj);
	hs_pending = true;  // rows and factors are rebuilt from this state at the next step
}

// The projection of an inequality row whose last build left the state s6
// (state columns 6-11) and geometry g: a limit, slide limit or motor row
// built beyond its maximum projects to <= 0, below its minimum to >= 0; a
// motor inside its range and a piston drive clamp to their bound and skip
// the position pass (TwistLimitConstraint / MotorConstraint /
// SlideLimitConstraint lim_status, ConstraintSolver.cpp:432-775).
int cp_world::state_projection(int kind, const double* s6, const double* g) {
	switch (kind) {
	case 9: return s6[0] == 0.0 && s6[1] > g[1] ? CP_PROJ_UPPER : CP_PROJ_LOWER;
	case 10: return s6[0] > g[1] ? CP_PROJ_UPPER : (s6[0] < g[0] ? CP_PROJ_LOWER : (CP_PROJ_CLAMP | CP_PROJ_NOPOS));
	case 11: return CP_PROJ_CLAMP | CP_PROJ_NOPOS;
	case 12: return s6[1] > g[10] ? CP_PROJ_UPPER : CP_PROJ_LOWER;
	default: return CP_PROJ_LOWER;
	}
}

// Static joint data, system membership, powered modes and the no-collision
// pair list; the accumulated state of joints already on the device is kept.
// New joints start as the reference's constructors leave them: limits
// inactive, and the positions of motors and pistons measured from the bodies
// at the next step (k_joint_init).
void cp_world::upload_joints() {
	int K = (int)j_type.size();
	const int K0 = std::max(j_dev_n, 0);
	std::vector<double> st(CP_JSTATE * (size_t)K, 0.0);
	// solver objects of the joints already on the device: rows, their
	// projections and pseudo impulses stay as they are (an inequality joint
	// may solve with them again); new joints start from zeroed rows
	std::vector<double> rows, pimp0;
	std::vector<int> proj0;
	if (K0 > 0) {
		HIPCHK(hipStreamSynchronize(stream));
		read_joint_state(st.data(), K0);
		down(rows, dj_row.p, (size_t)CP_JRS * K0);
		down(pimp0, dj_pimp.p, (size_t)CP_JMAX * K0);
		down(proj0, dj_proj.p, K0);
	}
	rows.resize((size_t)CP_JRS * std::max(K, 1), 0.0);
	for (int k = std::max(j_dev_n, 0); k < K; k++)
		if (j_kind[k] == 8 || j_kind[k] == 9 || j_kind[k] == 12) st[CP_JSTATE * (size_t)k + CP_JMAX] = 1.0;  // inactive
	std::vector<int> sys(K, -1);
	for (size_t q = 0; q < hs_og.size(); q++)
		for (int j : hs_og[q]) sys[j] = (int)q;
	up(dj_type, j_type); up(dj_kind, j_kind); up(dj_a, j_a); up(dj_b, j_b); up(dj_sys, sys);
	std::vector<double> g[4], ps(K), rs(K);
	for (auto& v : g) v.resize(3 * (size_t)K);
	for (int k = 0; k < K; k++) {
		const double* r = &j_geo[14 * (size_t)k];
		for (int v = 0; v < 4; v++)
			for (int c = 0; c < 3; c++) g[v][c * (size_t)K + k] = r[3 * v + c];
		ps[k] = r[12];
		rs[k] = r[13];
	}
	up(dj_tp1, g[0]); up(dj_tp2, g[1]); up(dj_ax1, g[2]); up(dj_ax2, g[3]); up(dj_ps, ps); up(dj_rs, rs);
	up(dj_cfm, j_cfm);
	upload_drives();
	up(dj_row, rows);
	j_dev_n = K;
	write_joint_state(st.data());
	if (K0 > 0) {
		std::vector<double> pimp(CP_JMAX * (size_t)K, 0.0);
		std::vector<int> proj;
		down(proj, dj_proj.p, K);
		for (int k = 0; k < K0; k++) {
			proj[k] = proj0[k];
			for (int c = 0; c < CP_JMAX; c++) pimp[c * (size_t)K + k] = pimp0[c * (size_t)K0 + k];
		}
		up(dj_pimp, pimp);
		up(dj_proj, proj);
	}
	// disallowCollision pairs of jointed bodies (PhysicsEngine.cpp:570-572),
	// sorted; springs leave their pair colliding (addSpring, :974-995)
	std::vector<uint64_t> nc;
	for (int k = 0; k < K; k++) {
		if (j_kind[k] == CP_JOINT_SPRING) continue;
		uint64_t lo = (uint64_t)std::min(j_a[k], j_b[k]), hi = (uint64_t)std::max(j_a[k], j_b[k]);
		nc.push_back(lo << 32 | hi);
	}
	std::sort(nc.begin(), nc.end());
	nc.erase(std::unique(nc.begin(), nc.end()), nc.end());
	up(nocol, nc);
	n_nocol = (int)nc.size();
	joints_dirty = false;
}

// Powered modes of motors and piston drives, component-major [4][K].
void cp_world::upload_drives() {
	const int K = (int)j_type.size();
	std::vector<double> d(4 * (size_t)std::max(K, 1), 0.0);
	for (int k = 0; k < K; k++)
		for (int c = 0; c < 4; c++) d[c * (size_t)K + k] = j_drive[4 * (size_t)k + c];
	up(dj_drive, d);
	drive_dirty = false;
}

// Motor angles and piston positions of joints added since the last step,
// measured from the current bodies (the Motor and Piston constructors).
void cp_world::init_new_joints(const Bodies& b) {
	const int K = (int)j_type.size();
	if (j_init_from >= K || !uploaded) return;
	launch_joint_init(b, joints(), j_init_from, K, stream);
	j_init_from = K;
}

// The order in which the reference applies spring impulses
// (PhysicsEngine.cpp:1313-1355): it walks the constraint graph's nodes in
// node-map order and each node's edges by the other body's id
// (insertNewEdge, :1984-1995), visiting an edge once, so a spring comes at
// (rank of its first-visited body, the other body), springs of one edge in
// creation order.  Springs that share no dynamic body commute; they are cut
// into groups (connected through shared dynamic bodies) that the device runs
// side by side, each in that order.
void cp_world::schedule_springs() {
	const int K = (int)j_type.size();
	const std::vector<int>& roots = root_order();
	if (sp_nj == K && sp_roots == roots) return;
	std::vector<int> sp;
	for (int k = 0; k < K; k++)
		if (j_kind[k] == CP_JOINT_SPRING) sp.push_back(k);
	std::vector<int> rank(nb, 0);
	for (int p = 0; p < (int)roots.size(); p++) rank[roots[p]] = p;
	auto key = [&](int k) {
		const int a = j_a[k], c = j_b[k];
		const int first = rank[a] < rank[c] ? a : c, other = first == a ? c : a;
		return std::make_pair(rank[first], other);
	};
	std::stable_sort(sp.begin(), sp.end(), [&](int x, int y) { return key(x) < key(y); });
	std::vector<int> parent(nb);
	for (int i = 0; i < nb; i++) parent[i] = i;
	auto find = [&](int i) {
		while (parent[i] != i) i = parent[i] = parent[parent[i]];
		return i;
	};
	for (int k : sp)
		if (b_move[j_a[k]] == 0 && b_move[j_b[k]] == 0) parent[find(j_a[k])] = find(j_b[k]);
	std::vector<int> group_of(nb, -1), gcount, gid(sp.size());
	for (size_t x = 0; x < sp.size(); x++) {
		const int k = sp[x];
		const int body = b_move[j_a[k]] == 0 ? j_a[k] : (b_move[j_b[k]] == 0 ? j_b[k] : -1);
		if (body < 0) {  // between two bodies that do not move: a group of its own
			gid[x] = (int)gcount.size();
			gcount.push_back(0);
		} else {
			int& g = group_of[find(body)];
			if (g < 0) {
				g = (int)gcount.size();
				gcount.push_back(0);
			}
			gid[x] = g;
		}
		gcount[gid[x]]++;
	}
	std::vector<int> goff(gcount.size() + 1, 0), list(sp.size());
	for (size_t g = 0; g < gcount.size(); g++) goff[g + 1] = goff[g] + gcount[g];
	std::vector<int> fill(goff.begin(), goff.end() - 1);
	for (size_t x = 0; x < sp.size(); x++) list[fill[gid[x]]++] = sp[x];
	up(sp_goff, goff);
	up(sp_list, list);
	sp_ngroups = (int)gcount.size();
	n_springs = (int)sp.size();
	sp_nj = K;
	sp_roots = roots;
}

// HolonomicSystem::HolonomicSystem (HolonomicBlockSolver.cpp:11-132): the
// elimination order that greedily adds the fewest fill-in blocks (ties to the
// first remaining constraint; a zero-fill candidate ends the search), then
// the block-lower-triangular layout of the factor in column-major order.
void cp_world::layout_systems() {
	hs_lay.assign(hs_og.size(), HLayout{});
	std::vector<int> nc, ord_off{0}, order, vloc, tab_off{0}, table, degree, buf_off{0}, dbuf_off{0};
	std::vector<int> cp_off, cs_ptr, cs_row, cs_loc, dg_loc;  // per column: the non-empty blocks below the diagonal
	for (size_t q = 0; q < hs_og.size(); q++) {
		const std::vector<int>& og = hs_og[q];
		HLayout& L = hs_lay[q];
		const int n = (int)og.size();
		auto share = [&](int x, int y) {
			int ax = j_a[og[x]], bx = j_b[og[x]], ay = j_a[og[y]], by = j_b[og[y]];
			return ax == ay || ax == by || bx == ay || bx == by;
		};
		std::vector<char> adj((size_t)n * n), filled((size_t)n * n);  // elimination graph; non-empty blocks
		for (int x = 0; x < n; x++)
			for (int y = 0; y < n; y++) adj[(size_t)x * n + y] = filled[(size_t)x * n + y] = share(x, y);
		std::vector<int> remaining(n), pos;
		for (int x = 0; x < n; x++) remaining[x] = x;
		while (!remaining.empty()) {
			int best = 0, best_fill = INT_MAX;
			std::vector<int> best_nb;
			for (int i = 0; i < (int)remaining.size(); i++) {
				int v = remaining[i];
				std::vector<int> nb;
				for (int y = 0; y < n; y++)
					if (y != v && adj[(size_t)v * n + y]) nb.push_back(y);
				int fill = 0;
				for (size_t x = 0; x < nb.size(); x++)
					for (size_t y = x + 1; y < nb.size(); y++) fill += !adj[(size_t)nb[x] * n + nb[y]];
				if (fill < best_fill) {
					best_fill = fill;
					best = i;
					best_nb.swap(nb);
					if (fill == 0) break;
				}
			}
			int v = remaining[best];
			pos.push_back(v);
			remaining.erase(remaining.begin() + best);
			for (int y = 0; y < n; y++) adj[(size_t)v * n + y] = adj[(size_t)y * n + v] = 0;
			for (size_t x = 0; x < best_nb.size(); x++)
				for (size_t y = x + 1; y < best_nb.size(); y++) {
					int a = best_nb[x], c = best_nb[y];
					adj[(size_t)a * n + c] = adj[(size_t)c * n + a] = 1;
					filled[(size_t)a * n + c] = filled[(size_t)c * n + a] = 1;
				}
		}
		for (int x : pos) L.order.push_back(og[x]);
		cp_off.push_back((int)cs_ptr.size());
		for (int col = 0; col < n; col++) {
			L.vloc.push_back(L.degree);
			L.degree += j_type[L.order[col]];
			cs_ptr.push_back((int)cs_row.size());
			for (int row = col; row < n; row++) {
				int l = -1;
				if (filled[(size_t)pos[row] * n + pos[col]]) {
					l = L.cap;
					L.cap += j_type[L.order[row]] * j_type[L.order[col]];
				}
				L.table.push_back(l);
				if (row == col) dg_loc.push_back(l);
				else if (l != -1) {
					cs_row.push_back(row);
					cs_loc.push_back(l);
				}
			}
		}
		cs_ptr.push_back((int)cs_row.size());
		nc.push_back(n);
		order.insert(order.end(), L.order.begin(), L.order.end());
		vloc.insert(vloc.end(), L.vloc.begin(), L.vloc.end());
		table.insert(table.end(), L.table.begin(), L.table.end());
		degree.push_back(L.degree);
		ord_off.push_back((int)order.size());
		tab_off.push_back((int)table.size());
		buf_off.push_back(buf_off.back() + L.cap);
		dbuf_off.push_back(dbuf_off.back() + 6 * L.degree);
	}
	up(dh_nc, nc); up(dh_ord_off, ord_off); up(dh_order, order); up(dh_vloc, vloc); up(dh_tab_off, tab_off); up(dh_table, table);
	up(dh_degree, degree); up(dh_buf_off, buf_off); up(dh_dbuf_off, dbuf_off);
	hs_lds_doubles = 0;
	for (size_t q = 0; q + 1 < buf_off.size(); q++)
		hs_lds_doubles = std::max(hs_lds_doubles, (buf_off[q + 1] - buf_off[q]) + (dbuf_off[q + 1] - dbuf_off[q]));
	if (cs_row.empty()) { cs_row.push_back(0); cs_loc.push_back(0); }
	if (cp_off.empty()) cp_off.push_back(0);
	if (cs_ptr.empty()) cs_ptr.push_back(0);
	if (dg_loc.empty()) dg_loc.push_back(0);
	up(dh_cp_off, cp_off); up(dh_cs_ptr, cs_ptr); up(dh_cs_row, cs_row); up(dh_cs_loc, cs_loc); up(dh_dg_loc, dg_loc);
	dh_buf.reserve(std::max(buf_off.back(), 1));
	dh_dbuf.reserve(std::max(dbuf_off.back(), 1));
	dh_scratch.reserve(std::max(dbuf_off.back() / 2, 1));
	HIPCHK(hipMemset(dh_buf.p, 0, std::max(buf_off.back(), 1) * sizeof(double)));
	hs_pending = true;
	joints_dirty = true;  // system membership changed
}

// StaticMeshGeometry::recomputeFromReference for every mesh body at its pose
// (RigidBody::updateGeometry, RigidBody.cpp:297-301): world triangles and the
// AABB tree built on the host, uploaded; the body's AABB is the mesh's.
void cp_world::place_meshes(const double* state13) {
	if (mesh_body.empty()) return;
	std::vector<double> tri, box;
	std::vector<int> tri_i, lr, root;
	mesh_depth = 0;
	std::vector<double> ab;
	down(ab, aabb.p, 6 * (size_t)nb);
	for (int body : mesh_body) {
		const double* r = state13 + 13 * (size_t)body;
		WorldMesh wm = place_mesh(hs.meshes[hs.mesh_of[b_shape[body]]], r, r + 3);
		int t0 = (int)(tri.size() / MT_REC), n0 = (int)(lr.size() / 2);
		tri.insert(tri.end(), wm.tri.begin(), wm.tri.end());
		tri_i.insert(tri_i.end(), wm.tri_i.begin(), wm.tri_i.end());
		box.insert(box.end(), wm.box.begin(), wm.box.end());
		for (size_t k = 0; k < wm.lr.size() / 2; k++) {
			bool leaf = wm.lr[2 * k] < 0;
			lr.push_back(leaf ? -1 : wm.lr[2 * k] + n0);
			lr.push_back(leaf ? wm.lr[2 * k + 1] + t0 : wm.lr[2 * k + 1] + n0);
		}
		root.push_back(wm.root + n0);
		mesh_depth = std::max(mesh_depth, wm.depth);
		for (int d = 0; d < 6; d++) ab[d * (size_t)nb + body] = wm.aabb[d];
	}
	if (mesh_depth + 1 > MESH_STACK) throw ApiErr{CP_ERR_INVALID};  // tree too deep for the device traversal
	up(dm_tri, tri); up(dm_tri_i, tri_i); up(dm_box, box); up(dm_lr, lr); up(dm_root, root);
	HIPCHK(hipMemcpy(aabb.p, ab.data(), ab.size() * sizeof(double), hipMemcpyHostToDevice));
	meshes_placed = true;
}

// Pairs with a static mesh: each takes MESH_SLOTS consecutive slots (same key)
// so that its several manifolds flow through persistence in order.
int cp_world::mesh_narrowphase(int npairs) {
	last_expanded = false;
	if (mesh_body.empty() || npairs <= 0) return npairs;
	if (!meshes_placed) throw ApiErr{CP_ERR_STATE};
	Bodies b = bodies();
	MeshDev md = meshdev();
	mesh_cnt.reserve(npairs + 1);
	mesh_off.reserve(npairs + 1);
	launch_mesh_expand(b, md, pairs.p, npairs, nocol.p, n_nocol, mesh_cnt.p, nullptr, nullptr, nullptr, stream);
	HIPCHK(hipMemsetAsync(mesh_cnt.p + npairs, 0, sizeof(int), stream));
	scan(mesh_cnt.p, mesh_off.p, npairs + 1);
	int total = read1(mesh_off.p + npairs);
	pairs2.reserve(total);
	pair_rank.reserve(total);
	launch_mesh_expand(b, md, pairs.p, npairs, nocol.p, n_nocol, mesh_cnt.p, mesh_off.p, pairs2.p, pair_rank.p, stream);
	std::swap(pairs, pairs2);
	last_expanded = true;
	return total;
}

// Body indices in the iteration order of the engine's node map (given by the
// caller, else libstdc++'s order for ids 1..n).
const std::vector<int>& cp_world::root_order() {
	if ((int)node_order.size() == nb) {
		root_cache.clear();
		for (int64_t id : node_order) root_cache.push_back((int)id - 1);
		root_cache_nb = -1;
	} else if (root_cache_nb != nb) {
		std::vector<int64_t> ord(nb);
		cp_reference_node_order(nb, ord.data());
		root_cache.assign(ord.begin(), ord.end());
		for (int& r : root_cache) r -= 1;
		root_cache_nb = nb;
	}
	return root_cache;
}

// maintainHolonomicSystems (PhysicsEngine.cpp:1292-1307) for joints added
// since the last step: in node-map order, each node's edges by the other
// body's id; an edge whose scan is pending collects the edges connected to it
// through holonomic joints (getAllEdgesConnectedHolonomically, :1746-1779:
// depth first, through dynamic bodies only) and, when there are at least two,
// founds a system or merges the systems met (:1466-1512).  Where the
// reference keeps the system that sorts first by address, the lowest system
// index is kept.  A system's constraint order is its member edges' joints
// (revaluateSystem, :1785-1798).
void cp_world::discover_systems() {
	std::vector<std::vector<std::pair<int, uint64_t>>> adj(nb);
	for (auto& kv : jedges) {
		adj[kv.second.n1].push_back({kv.second.n2, kv.first});
		adj[kv.second.n2].push_back({kv.second.n1, kv.first});
	}
	for (auto& a : adj) std::sort(a.begin(), a.end());
	bool changed = hs_changed;
	for (int node : root_order()) {
		for (auto& ne : adj[node]) {
			JEdge& e = jedges[ne.second];
			if (!e.scan) continue;
			e.scan = false;
			std::vector<uint64_t> out, st{ne.second};
			std::vector<uint64_t> visited;
			auto seen = [&](uint64_t k) { return std::find(visited.begin(), visited.end(), k) != visited.end(); };
			while (!st.empty()) {
				uint64_t k = st.back();
				st.pop_back();
				if (seen(k)) continue;
				visited.push_back(k);
				out.push_back(k);
				const JEdge& v = jedges[k];
				std::vector<uint64_t> nbrs;
				if (b_move[v.n1] == 0)
					for (auto& x : adj[v.n1])
						if (jedges[x.second].nhol) nbrs.push_back(x.second);
				if (b_move[v.n2] == 0)
					for (auto& x : adj[v.n2])
						if (jedges[x.second].nhol) nbrs.push_back(x.second);
				for (uint64_t x : nbrs)
					if (!seen(x)) st.push_back(x);
			}
			if (out.size() <= 1) continue;
			int rep = INT_MAX;
			bool any_null = false;
			std::vector<int> met;
			for (uint64_t k : out) {
				int q = jedges[k].sys;
				if (q < 0) any_null = true;
				else {
					rep = std::min(rep, q);
					if (std::find(met.begin(), met.end(), q) == met.end()) met.push_back(q);
				}
			}
			if (met.size() == 1 && !any_null) continue;  // already one system, nothing new
			if (met.empty()) {
				rep = (int)sys_edges.size();
				sys_edges.emplace_back();
			}
			for (uint64_t k : out) {
				JEdge& x = jedges[k];
				if (x.sys != rep) sys_edges[rep].push_back(k);
				x.scan = false;
			}
			for (uint64_t k : out) jedges[k].sys = rep;
			// merged systems are dropped; renumber the rest
			std::vector<std::vector<uint64_t>> keep;
			for (int q = 0; q < (int)sys_edges.size(); q++)
				if (q == rep || std::find(met.begin(), met.end(), q) == met.end()) keep.push_back(sys_edges[q]);
			sys_edges.swap(keep);
			for (int q = 0; q < (int)sys_edges.size(); q++)
				for (uint64_t k : sys_edges[q]) jedges[k].sys = q;
			changed = true;
		}
	}
	hs_changed = false;
	if (!changed) return;
	hs_og.clear();
	for (auto& edges : sys_edges) {
		hs_og.emplace_back();
		for (uint64_t k : edges)
			for (int j : jedges[k].joints)
				if (!joint_is_inequality(j_kind[j])) hs_og.back().push_back(j);  // revaluateSystem (:1785-1798)
	}
	layout_systems();
}

// ---------------------------------------------------------------- serving
// Shape tables are append-only on the host; only their new tails go to the device.
void cp_world::sync_shapes() {
	HostShapes& h = hs;
	int k = 0;
	auto ints = [&](DArr<int>& d, std::vector<int>& v) { up_tail(d, v, sh_done[k++]); };
	auto dbls = [&](DArr<double>& d, std::vector<double>& v) { up_tail(d, v, sh_done[k++]); };
	ints(s_type, h.type); ints(s_pt_off, h.pt_off); ints(s_face_off, h.face_off); ints(s_arc_off, h.arc_off); ints(s_edge_off, h.edge_off);
	dbls(s_ref_pts, h.ref_pts); ints(s_fv_off, h.fv_off); ints(s_fv, h.fv); ints(s_ncalc, h.ncalc); ints(s_ndir, h.ndir); ints(s_sid, h.sid);
	dbls(s_gnorm, h.gnorm); ints(s_red, h.red); ints(s_minpid, h.minpid); ints(s_maxpid, h.maxpid); ints(s_refidx, h.refidx);
	dbls(s_gmin, h.gmin); dbls(s_gmax, h.gmax); dbls(s_grefval, h.grefval); ints(s_arcs, h.arcs); ints(s_edges, h.edges);
	ints(s_vaf_off, h.vaf_off); ints(s_vaf, h.vaf); ints(s_vae_off, h.vae_off); ints(s_vae, h.vae);
	dbls(s_radius, h.radius); dbls(s_height, h.height);
}

void cp_world::stage_body(int i, int shape_id, int move, const double* mp19, const double* mat4, const double* st13) {
	b_shape[i] = shape_id;
	b_move[i] = move;
	for (int c = 0; c < 19; c++) b_massprops[19 * (size_t)i + c] = mp19[c];
	for (int c = 0; c < 4; c++) b_mat[4 * (size_t)i + c] = mat4[c];
	stage_i.insert(stage_i.end(), {i, shape_id, move});
	stage_d.insert(stage_d.end(), mp19, mp19 + 19);
	stage_d.insert(stage_d.end(), mat4 + 1, mat4 + 4);
	stage_d.insert(stage_d.end(), st13, st13 + 13);
}

// Parked bodies: FIXED unit spheres far from everything, 3 apart.
void cp_world::park_range(int first, int count) {
	double mp[19] = {1, 1, 0, 0, 0, 1, 0, 0, 0, 1, 1, 0, 0, 0, 1, 0, 0, 0, 1};
	double mat4[4] = {0, 0, 0, 0};
	for (int i = first; i < first + count; i++) {
		double st[13] = {1.0e7 + 3.0 * i, -1.0e7, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0};
		stage_body(i, park_shape, 1, mp, mat4, st);
	}
}

// Applies staged admissions and retirements: new shape-table tails, one body
// scatter, contact retirement, scene offsets and size-class lists.
void cp_world::serving_commit() {
	if (!serving_dirty) return;
	sync_shapes();
	int nrec = (int)stage_i.size() / 3;
	if (nrec) {
		// a body staged more than once (retired, then admitted again before this
		// commit) keeps its last record: the scatter writes each body once
		if ((int)stage_mark.size() < nb) stage_mark.assign(nb, -1);
		commit_id++;
		std::vector<int> ri, list;
		std::vector<double> rd;
		for (int k = nrec - 1; k >= 0; k--) {
			int i = stage_i[3 * k];
			if (stage_mark[i] == commit_id) continue;
			stage_mark[i] = commit_id;
			ri.insert(ri.end(), stage_i.begin() + 3 * k, stage_i.begin() + 3 * k + 3);
			rd.insert(rd.end(), stage_d.begin() + (size_t)BODY_REC * k, stage_d.begin() + (size_t)BODY_REC * (k + 1));
			list.push_back(i);
		}
		up(st_i, ri);
		up(st_d, rd);
		up(st_list, list);
		launch_body_scatter(bodies(), shapes(), (int)list.size(), st_i.p, st_d.p, st_list.p, stream);
		launch_zero_list(nf_body.p, st_list.p, (int)list.size(), stream);  // admitted bodies start unflagged
	}
	if (!kill_ranges.empty()) {
		up(st_kill, kill_ranges);
		launch_kill_contacts(cb[cur].c, ncontacts, st_kill.p, (int)kill_ranges.size() / 2, stream);
	}
	apply_resumes();
	stage_i.clear();
	stage_d.clear();
	kill_ranges.clear();
	// every range is a scene (free ranges hold parked bodies and never pair);
	// only admitted scenes enter the solver's size-class lists
	std::vector<int> sbo;
	for (const Range& r : ranges) sbo.push_back(r.start);
	sbo.push_back(nb);
	nscenes = (int)ranges.size();
	up(scene_body_off, sbo);
	scene_of_body.reserve(nb);
	launch_scene_of_body(nb, scene_body_off.p, nscenes, scene_of_body.p, stream);
	live_n = 0;
	live_scenes = 0;
	for (int q = 0; q < nscenes; q++)
		if (ranges[q].handle >= 0) {
			live_n = ranges[q].start + ranges[q].len;
			live_scenes = q + 1;
		}
	cls_sizes.clear();
	for (int q = 0; q < nscenes; q++)
		if (ranges[q].handle >= 0) cls_sizes.push_back({q, ranges[q].len});
	build_classes();
	HIPCHK(hipStreamSynchronize(stream));
	serving_dirty = false;
}

// PGS size classes: scenes whose velocity deltas fit in LDS run the LDS
// kernel of their class (largest first within a class); larger scenes use
// global memory.
void cp_world::build_classes() {
	const int* lim = cfg.pgs_class_bodies;
	cls_lists.assign(4, std::vector<int>());
	std::vector<int> size_of(nscenes + 1, 0);
	for (auto& q : cls_sizes) {
		int m = q.second;
		cls_lists[m <= lim[0] ? 0 : m <= lim[1] ? 1 : m <= lim[2] ? 2 : 3].push_back(q.first);
		size_of[q.first] = m;
	}
	for (int c = 0; c < 4; c++) {
		std::sort(cls_lists[c].begin(), cls_lists[c].end(), [&](int a, int b2) { return size_of[a] > size_of[b2]; });
		cls_max[c] = cls_lists[c].empty() ? 0 : size_of[cls_lists[c][0]];
		up(cls_dev[c], cls_lists[c]);
	}
	for (int c = 0; c < 3; c++) cls_built[c] = lim[c];
	max_scene_bodies = 0;
	for (auto& q : cls_sizes) max_scene_bodies = std::max(max_scene_bodies, q.second);
}

void cp_world::scan(const int* in, int* out, int n) {
	size_t bytes = 0;
	HIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, n, stream));
	tmp.reserve(bytes + 16);
	HIPCHK(hipcub::DeviceScan::ExclusiveSum(tmp.p, bytes, in, out, n, stream));
}

void cp_world::ensure_manifolds(int n) {
	if (n <= mcap) return;
	for (void* a : man_allocs) (void)hipFree(a);
	man_allocs.clear();
	int K = std::max(n + n / 2, 1024);
	auto al = [&](size_t bytes) {
		void* p = nullptr;
		if (hipMalloc(&p, bytes) != hipSuccess) throw ApiErr{CP_ERR_OOM};
		man_allocs.push_back(p);
		return p;
	};
	man.count = (int*)al(sizeof(int) * K);
	man.normal = (double*)al(sizeof(double) * 3 * K);
	man.maxpen = (double*)al(sizeof(double) * K);
	man.pts = (double*)al(sizeof(double) * 16 * K);
	man.magic = (uint64_t*)al(sizeof(uint64_t) * 8 * K);
	man.mat = (double*)al(sizeof(double) * 12 * K);
	mcap = K;
}

void cp_world::ensure_points(int n) {
	if (n <= npcap) return;
	for (void* a : np_allocs) (void)hipFree(a);
	np_allocs.clear();
	int K = std::max(n + n / 2, 1024);
	auto al = [&](size_t bytes) {
		void* p = nullptr;
		if (hipMalloc(&p, bytes) != hipSuccess) throw ApiErr{CP_ERR_OOM};
		np_allocs.push_back(p);
		return p;
	};
	np.key = (uint64_t*)al(8 * (size_t)K); np.cid = (uint64_t*)al(8 * (size_t)K); np.bid = (uint64_t*)al(8 * (size_t)K);
	np.pos = (double*)al(24 * (size_t)K); np.nrm = (double*)al(24 * (size_t)K);
	np.pen = (double*)al(8 * (size_t)K); np.mat = (double*)al(24 * (size_t)K);
	np.is_new = (int*)al(4 * (size_t)K); np.src = (int*)al(4 * (size_t)K);
	npcap = K;
}

// Returns the number of candidate pairs now in `pairs` (sorted keys).
int cp_world::broadphase() {
	Bodies b = bodies();
	int n = b.nl;
	int ns = serving ? live_scenes : nscenes;
	if (n == 0) return 0;
	bp_keys.reserve(n); bp_keys2.reserve(n); bp_vals.reserve(n); bp_vals2.reserve(n);
	bp_sb.reserve(6 * (size_t)n);
	bp_counts.reserve(n + 1);
	launch_bp_keys(b, bp_keys.p, bp_vals.p, stream);
	size_t bytes = 0;
	if (ns == 1) {
		HIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, bp_keys.p, bp_keys2.p, bp_vals.p, bp_vals2.p, n, 0, 64, stream));
		tmp.reserve(bytes + 16);
		HIPCHK(hipcub::DeviceRadixSort::SortPairs(tmp.p, bytes, bp_keys.p, bp_keys2.p, bp_vals.p, bp_vals2.p, n, 0, 64, stream));
	} else {
		HIPCHK(hipcub::DeviceSegmentedRadixSort::SortPairs(nullptr, bytes, bp_keys.p, bp_keys2.p, bp_vals.p, bp_vals2.p, n, ns,
		                                                   scene_body_off.p, scene_body_off.p + 1, 0, 64, stream));
		tmp.reserve(bytes + 16);
		HIPCHK(hipcub::DeviceSegmentedRadixSort::SortPairs(tmp.p, bytes, bp_keys.p, bp_keys2.p, bp_vals.p, bp_vals2.p, n, ns,
		                                                   scene_body_off.p, scene_body_off.p + 1, 0, 64, stream));
	}
	launch_bp_gather(b, bp_vals2.p, bp_sb.p, (cfg.defect_fixes >> 1) & 1, stream);
	// sweep ends and chunks, then count / scan / write per chunk
	bp_ext.reserve(n); bp_choff.reserve(n + 1);
	HIPCHK(hipMemsetAsync(bp_counts.p + n, 0, sizeof(int), stream));
	launch_bp_extent(n, bp_vals2.p, bp_sb.p, scene_of_body.p, scene_body_off.p, bp_ext.p, bp_counts.p, stream);
	scan(bp_counts.p, bp_choff.p, n + 1);
	int nch = read1(bp_choff.p + n);
	bp_owner.reserve(nch); bp_ccount.reserve(nch + 1); bp_coffs.reserve(nch + 1);
	launch_bp_owner(n, bp_choff.p, bp_owner.p, stream);
	HIPCHK(hipMemsetAsync(bp_ccount.p + nch, 0, sizeof(int), stream));
	launch_bp_sweep(n, nch, bp_vals2.p, bp_sb.p, bp_ext.p, bp_choff.p, bp_owner.p, bp_ccount.p, nullptr, nullptr, 0, stream);
	scan(bp_ccount.p, bp_coffs.p, nch + 1);
	int np_ = read1(bp_coffs.p + nch);
	pairs.reserve(np_); pairs2.reserve(np_);
	launch_bp_sweep(n, nch, bp_vals2.p, bp_sb.p, bp_ext.p, bp_choff.p, bp_owner.p, bp_ccount.p, bp_coffs.p, pairs2.p, 1, stream);
	if (np_ > 0) {
		bytes = 0;
		HIPCHK(hipcub::DeviceRadixSort::SortKeys(nullptr, bytes, pairs2.p, pairs.p, np_, 0, 64, stream));
		tmp.reserve(bytes + 16);
		HIPCHK(hipcub::DeviceRadixSort::SortKeys(tmp.p, bytes, pairs2.p, pairs.p, np_, 0, 64, stream));
	}
	return np_;
}

int cp_world::narrowphase(int npairs) {
	npairs = mesh_narrowphase(npairs);
	ensure_manifolds(npairs);
	np_lists.reserve(7 * (size_t)std::max(npairs, 1));
	np_counts.reserve(8);
	launch_narrowphase(bodies(), shapes(), params(), pairs.p, npairs, man, mcap, np_lists.p, np_counts.p, nocol.p, n_nocol,
	                   last_expanded ? pair_rank.p : nullptr, stream, &fk);
	if (!last_expanded) return npairs;
	// convex body against mesh: candidate triangles, per-triangle SAT, merge + cull
	int nlist = read1(np_counts.p + NPC_MESH_LIST);
	if (nlist == 0) return npairs;
	const int* list = np_lists.p + (size_t)NPC_MESH_LIST * npairs;
	Bodies b = bodies();
	MeshDev md = meshdev();
	mesh_ccnt.reserve(nlist + 1);
	mesh_coff.reserve(nlist + 1);
	mesh_overflow.reserve(2);
	HIPCHK(hipMemsetAsync(mesh_overflow.p, 0, 2 * sizeof(int), stream));
	MeshItems it = meshitems();
	launch_mesh_candidates(b, md, pairs.p, list, np_counts.p, nlist, mesh_ccnt.p, nullptr, it, mesh_overflow.p, stream);
	HIPCHK(hipMemsetAsync(mesh_ccnt.p + nlist, 0, sizeof(int), stream));
	scan(mesh_ccnt.p, mesh_coff.p, nlist + 1);
	int C = read1(mesh_coff.p + nlist);
	size_t c = std::max(C, 1);
	mi_slot.reserve(c); mi_tri.reserve(c); mi_count.reserve(c); mi_owner.reserve(c); mi_avail.reserve(c);
	mi_head.reserve(4 * c); mi_pts.reserve(4 * MESH_PTS * c); mi_magic.reserve(2 * MESH_PTS * c); mi_mat.reserve(3 * c);
	it = meshitems();
	launch_mesh_candidates(b, md, pairs.p, list, np_counts.p, nlist, mesh_ccnt.p, mesh_coff.p, it, mesh_overflow.p, stream);
	launch_mesh_triangles(b, shapes(), md, params(), pairs.p, it, C, stream);
	launch_mesh_merge(params(), pairs.p, md, list, nlist, mesh_coff.p, it, man, mcap, mesh_overflow.p + 1, stream);
	int ov[2];
	HIPCHK(hipMemcpyAsync(ov, mesh_overflow.p, sizeof(ov), hipMemcpyDeviceToHost, stream));
	HIPCHK(hipStreamSynchronize(stream));
	if (ov[0] || ov[1]) throw ApiErr{CP_ERR_CAPACITY};  // deeper tree or more manifolds per pair than the slots hold
	return npairs;
}

// Contact persistence: compaction of manifold points, match, merge.
void cp_world::persist(int npairs) {
	Bodies b = bodies();
	Params P = params();
	pt_off.reserve(npairs + 1);
	pt_cnt.reserve(npairs + 1);
	if (npairs) HIPCHK(hipMemcpyAsync(pt_cnt.p, man.count, npairs * sizeof(int), hipMemcpyDeviceToDevice, stream));
	HIPCHK(hipMemsetAsync(pt_cnt.p + npairs, 0, sizeof(int), stream));
	scan(pt_cnt.p, pt_off.p, npairs + 1);
	int nnew = read1(pt_off.p + npairs);
	live_contacts = nnew;
	ensure_points(nnew);
	launch_np_gather(man, mcap, pairs.p, npairs, pt_off.p, np, npcap, stream);
	int nold = ncontacts;
	if (nold == 0) cb[cur].reserve(1);
	launch_np_match(np, nnew, npcap, cb[cur].c, nold, b, P, stream);
	keep.reserve(nold + 1); keep_scan.reserve(nold + 1);
	launch_old_keep(cb[cur].c, nold, keep.p, b.asleep_rm, stream);
	scan(keep.p, keep_scan.p, nold + 1);
	// is_new -> rank (exclusive scan over nnew+1 with a zero tail)
	rank.reserve(nnew + 1);
	pt_cnt.reserve(nnew + 1);
	if (nnew) HIPCHK(hipMemcpyAsync(pt_cnt.p, np.is_new, nnew * sizeof(int), hipMemcpyDeviceToDevice, stream));
	HIPCHK(hipMemsetAsync(pt_cnt.p + nnew, 0, sizeof(int), stream));
	scan(pt_cnt.p, rank.p, nnew + 1);
	int nkept = read1(keep_scan.p + nold);
	int nnc = read1(rank.p + nnew);
	nk.reserve(nnc + 1);
	launch_new_keys(np, nnew, rank.p, nk.p, stream);
	int other = 1 - cur;
	int total = nkept + nnc;
	cb[other].reserve(total);
	launch_merge(cb[cur].c, nold, keep_scan.p, nk.p, nnc, np, nnew, npcap, rank.p, cb[other].c, b, P, seq, stream);
	seq += (uint64_t)nnew + 1;
	cur = other;
	ncontacts = total;
}

void cp_world::edges() {
	int n = ncontacts;
	e_head.reserve(n + 1); e_idx.reserve(n + 1);
	HIPCHK(hipMemsetAsync(e_head.p + n, 0, sizeof(int), stream));
	launch_edge_heads(cb[cur].c, n, e_head.p, stream);
	scan(e_head.p, e_idx.p, n + 1);
	nedges = read1(e_idx.p + n);
	e_key.reserve(nedges); e_first.reserve(nedges); e_cnt.reserve(nedges); e_live.reserve(nedges);
	e_color.reserve(nedges); e_order.reserve(nedges);
	launch_edge_build(cb[cur].c, n, e_head.p, e_idx.p, e_key.p, e_first.p, e_cnt.p, e_live.p, stream);
}

// Island search of PhysicsEngine::sleepOrSolveIslands (PhysicsEngine.cpp:1536-1613)
// on the device (k_islands.hip) over the post-narrowphase cache and the
// joints.  The host supplies the joint pair groups and the node-map order
// when they change and reads back the island count and the largest island.
// Edges are body pairs with contacts or joints, a node's edges ordered by
// the other body's id (ConstraintGraphNode::insertNewEdge, :1984-1995); an
// edge's rows are its joints in creation order, then its live contacts.
void cp_world::schedule_islands(bool colored) {
	const int n = ncontacts, nj = (int)j_type.size();
	if (ib_nj != nj) {
		// joint pair groups in key order (lo << 32 | hi), joints of a group in
		// creation order
		std::vector<uint64_t> jkey(nj), gk;
		std::vector<int> jord(nj), goff;
		for (int j = 0; j < nj; j++) {
			uint64_t lo = (uint64_t)std::min(j_a[j], j_b[j]), hi = (uint64_t)std::max(j_a[j], j_b[j]);
			jkey[j] = lo << 32 | hi;
			jord[j] = j;
		}
		std::stable_sort(jord.begin(), jord.end(), [&](int x, int y) { return jkey[x] < jkey[y]; });
		for (int x = 0; x < nj; x++)
			if (x == 0 || jkey[jord[x]] != jkey[jord[x - 1]]) {
				gk.push_back(jkey[jord[x]]);
				goff.push_back(x);
			}
		goff.push_back(nj);
		up(ib_jord, jord);
		up(ib_jgkey, gk);
		up(ib_jgoff, goff);
		ib_ng = (int)gk.size();
		ib_nj = nj;
	}
	const std::vector<int>& roots = root_order();
	if (ib_rord_host != roots) {
		ib_rord_host = roots;
		std::vector<int> rank(nb, 0);
		for (int p = 0; p < (int)roots.size(); p++) rank[roots[p]] = p;
		up(ib_rord, roots);
		up(ib_rank, rank);
	}
	const int ne = nedges, ng = ib_ng, N = ne + ng;
	const size_t NB = (size_t)nb;
	ib_keys.reserve(N); ib_keys2.reserve(N); ib_vals.reserve(N); ib_vals2.reserve(N);
	ib_head.reserve(N + 1); ib_midx.reserve(N + 1); ib_me.reserve(8 * (size_t)N);
	ib_adjk.reserve(2 * (size_t)N); ib_adjk2.reserve(2 * (size_t)N); ib_adjv.reserve(2 * (size_t)N); ib_adjv2.reserve(2 * (size_t)N);
	ib_aoff.reserve(NB + 1); ib_parent.reserve(NB); ib_comp.reserve(NB);
	ib_acc.reserve(NB); ib_in.reserve(NB + 1); ib_out.reserve(NB + 1); ib_info.reserve(2);
	DArr<int>* per_isl[] = {&isl_root, &isl_edge_off, &isl_body_off, &isl_body_cnt, &isl_stack_off, &isl_joint_off, &isl_icoff, &isl_ncol,
	                        &isl_serial, &isl_nbod, &isl_sys_cnt};
	for (auto* a : per_isl) a->reserve(NB + 1);
	isl_off.reserve(NB + 1);
	ib_stack.reserve(2 * (size_t)N + NB); ib_vlist.reserve(NB); ib_elist.reserve(N); ib_ecol.reserve(N); ib_eorder.reserve(N);
	isl_eoff.reserve(N + 1); isl_coff.reserve((size_t)N + NB + 1); isl_bod.reserve(NB);
	isl_rows.reserve(2 * ((size_t)nj + 2 * (size_t)n)); isl_cloc.reserve(2 * (size_t)n); isl_jloc.reserve(2 * (size_t)nj);
	isl_sys.reserve(nj); ib_slots.reserve(2 * (size_t)nj); isl_jlin.reserve(2 * (size_t)nj);
	// walk scratch that every walk leaves as it found it
	auto keep_filled = [&](auto& a, size_t m, int byte) {
		if (a.p && m <= a.cap) return;
		a.reserve(m);
		HIPCHK(hipMemsetAsync(a.p, byte, a.cap * sizeof(*a.p), stream));
	};
	keep_filled(ib_vis, NB, 0);
	keep_filled(ib_loc, NB, 0xFF);
	keep_filled(ib_col_used, 2 * NB, 0);
	keep_filled(ib_sys_seen, hs_lay.size() + 1, 0);
	size_t bytes = 0;
	if (N) {
		launch_ib_keys(e_key.p, ne, ib_jgkey.p, ng, ib_keys.p, ib_vals.p, stream);
		HIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, ib_keys.p, ib_keys2.p, ib_vals.p, ib_vals2.p, N, 0, 64, stream));
		tmp.reserve(bytes + 16);
		HIPCHK(hipcub::DeviceRadixSort::SortPairs(tmp.p, bytes, ib_keys.p, ib_keys2.p, ib_vals.p, ib_vals2.p, N, 0, 64, stream));
	}
	launch_ib_heads(ib_keys2.p, N, ib_head.p, stream);
	scan(ib_head.p, ib_midx.p, N + 1);
	if (N) {
		launch_ib_edges(ib_keys2.p, ib_vals2.p, N, ib_head.p, ib_midx.p, ne, e_first.p, e_cnt.p, cb[cur].c.live, ib_jgoff.p, ib_jord.p,
		                joint_act(), ib_me.p, ib_adjk.p, ib_adjv.p, stream);
		bytes = 0;
		HIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, ib_adjk.p, ib_adjk2.p, ib_adjv.p, ib_adjv2.p, 2 * N, 0, 64, stream));
		tmp.reserve(bytes + 16);
		HIPCHK(hipcub::DeviceRadixSort::SortPairs(tmp.p, bytes, ib_adjk.p, ib_adjk2.p, ib_adjv.p, ib_adjv2.p, 2 * N, 0, 64, stream));
	}
	launch_ib_aoff(ib_adjk2.p, 2 * N, nb, ib_aoff.p, stream);
	const SleepDev S = sleep_dev();
	launch_ib_components(ib_me.p, ib_midx.p, N, nb, movement.p, ib_rank.p, ib_aoff.p, ib_parent.p, ib_comp.p, ib_acc.p, S, stream);
	launch_ib_flags(nb, ib_rord.p, ib_comp.p, ib_acc.p, S.sleeping, ib_in.p, stream);
	bytes = 0;
	const IslCnt zero{0, 0, 0, 0, 0, 0, 0};
	HIPCHK(hipcub::DeviceScan::ExclusiveScan(nullptr, bytes, ib_in.p, ib_out.p, IslCntSum(), zero, nb + 1, stream));
	tmp.reserve(bytes + 16);
	HIPCHK(hipcub::DeviceScan::ExclusiveScan(tmp.p, bytes, ib_in.p, ib_out.p, IslCntSum(), zero, nb + 1, stream));
	HIPCHK(hipMemsetAsync(ib_info.p, 0, 2 * sizeof(int), stream));
	const IslTables T = isl_tables();
	launch_ib_tables(nb, ib_rord.p, ib_in.p, ib_out.p, T, ib_info.p, ib_comp.p, ib_acc.p, S, vel.p, angv.p, stream);
	int info[2];
	HIPCHK(hipMemcpyAsync(info, ib_info.p, 2 * sizeof(int), hipMemcpyDeviceToHost, stream));
	HIPCHK(hipStreamSynchronize(stream));
	isl_n = info[0];
	IslWalk W{};
	W.T = T;
	W.nb = nb;
	W.hol = cfg.hol_iters > 0 && !hs_lay.empty();
	W.movement = movement.p; W.asleep = S.enabled ? S.asleep : nullptr; W.aoff = ib_aoff.p; W.adjv = ib_adjv2.p; W.me = ib_me.p;
	W.jord = ib_jord.p; W.live = cb[cur].c.live; W.ja = dj_a.p; W.jb = dj_b.p; W.jsys = dj_sys.p; W.adjk = ib_adjk2.p;
	W.jact = joint_act();
	W.ckey = cb[cur].c.key; W.stack = ib_stack.p; W.vlist = ib_vlist.p; W.elist = ib_elist.p; W.vis = ib_vis.p; W.rows = isl_rows.p;
	W.ecol = ib_ecol.p; W.coff = isl_coff.p; W.eorder = ib_eorder.p; W.icoff = isl_icoff.p; W.ncol = isl_ncol.p; W.serial = isl_serial.p;
	W.eoff = isl_eoff.p; W.bod = isl_bod.p; W.nbod = isl_nbod.p; W.loc = ib_loc.p; W.cloc = isl_cloc.p; W.jloc = isl_jloc.p;
	W.sys_seen = ib_sys_seen.p; W.sys = isl_sys.p; W.slots = ib_slots.p; W.jlin = isl_jlin.p; W.sys_cnt = isl_sys_cnt.p;
	W.col_used = ib_col_used.p;
	launch_ib_walk(W, isl_n, colored, stream);
	lin_nl = nb + 2 * nj;
	isl_lds_bodies = std::min(info[1] + 1, ISL_LDS_BODIES);
	islands_on_device = true;
}

SleepDev cp_world::sleep_dev() {
	SleepDev S{};
	S.enabled = sl_active ? 1 : 0;
	S.sleeping = sl_active && cfg.sleeping ? 1 : 0;
	if (!S.enabled) return S;
	S.asleep = sl_asleep.p; S.counter = sl_counter.p; S.ticks = sl_ticks.p; S.hlen = sl_hlen.p; S.hhead = sl_hhead.p;
	S.hist = sl_hist.p; S.disabled = sl_disabled.p; S.proc = nullptr;
	S.hcap = sl_hcap;
	// PhysicsEngine::bodySleepy (PhysicsEngine.cpp:1243-1245) and the wake
	// threshold of applyVelocityChange (:1124), as the reference computes them
	const double* g = cfg.gravity;
	const double gmag = std::sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
	S.delay = cfg.sleep_delay;
	S.step_time = cfg.dt;
	S.vel_eps = cfg.sleep_vel_coeff * gmag;
	S.accel_eps = cfg.sleep_accel_coeff * gmag;
	S.ang_accel_eps = cfg.sleep_accel_coeff * 2.5;
	S.ang_vel_eps = cfg.sleep_ang_vel_eps;
	S.wake_vel = 0.25 * cfg.sleep_vel_coeff * gmag;
	S.hist_need = cfg.sleep_delay / cfg.dt;
	S.tick_threshold = cfg.sleep_tick_threshold;
	return S;
}

// Sleep state at the step start: the history ring sized for
// (int)(sleep_delay / dt) entries (RigidBody::recordMovementState's length,
// PhysicsEngine.cpp:72), and the waking actions since the last step
// (PhysicsEngine.cpp:66-69): in body order, a body with a waking action wakes
// every body of its island in the constraint graph as it stands (the contact
// cache without the contacts the last step removed, and the joints); a body
// the loop reaches while still asleep gets no history entry and no gravity
// this step, even when a later body's action wakes it.  proc is set when
// those per-body decisions were uploaded.
void cp_world::prepare_sleep(bool& proc) {
	proc = false;
	const int hcap = static_cast<int>(cfg.sleep_delay / cfg.dt);
	if (hcap != sl_hcap) {
		sl_hist.reserve(6 * (size_t)std::max(hcap, 1) * nb);
		HIPCHK(hipMemsetAsync(sl_hlen.p, 0, (size_t)nb * sizeof(int), stream));
		HIPCHK(hipMemsetAsync(sl_hhead.p, 0, (size_t)nb * sizeof(int), stream));
		sl_hcap = hcap;
	}
	if (sl_wake.empty()) return;
	HIPCHK(hipStreamSynchronize(stream));
	std::vector<int> asleep, rm, ticks, hlen, hhead, live;
	std::vector<double> counter;
	std::vector<int> life;
	std::vector<uint64_t> key;
	down(asleep, sl_asleep.p, nb); down(rm, sl_asleep_rm.p, nb); down(ticks, sl_ticks.p, nb);
	down(hlen, sl_hlen.p, nb); down(hhead, sl_hhead.p, nb); down(counter, sl_counter.p, nb);
	down(key, cb[cur].c.key, ncontacts); down(life, cb[cur].c.life, ncontacts);
	std::vector<std::vector<int>> adj(nb);
	for (int k = 0; k < ncontacts; k++) {
		int a = (int)(key[k] >> 32), c = (int)(key[k] & 0xFFFFFFFFu);
		if (life[k] < -1 && !rm[a] && !rm[c]) continue;  // removed at the last step's first substep
		adj[a].push_back(c);
		adj[c].push_back(a);
	}
	for (size_t j = 0; j < j_type.size(); j++) {
		adj[j_a[j]].push_back(j_b[j]);
		adj[j_b[j]].push_back(j_a[j]);
	}
	std::vector<char> req(nb, 0), vis(nb, 0);
	for (int i : sl_wake) req[i] = 1;
	sl_wake.clear();
	std::vector<int> pr(nb, 0), st;
	for (int i = 0; i < nb; i++) {
		if (req[i]) {
			// wakeupIsland (PhysicsEngine.cpp:1285-1290) through bfsVisitAll (:1263-1283)
			std::fill(vis.begin(), vis.end(), 0);
			st.assign(1, i);
			while (!st.empty()) {
				int c = st.back();
				st.pop_back();
				if (vis[c]) continue;
				vis[c] = 1;
				if (b_move[c] == 0 && asleep[c]) {
					asleep[c] = 0;
					counter[c] = 0;
					ticks[c] = 0;
					hlen[c] = 0;
					hhead[c] = 0;
				}
				for (int o : adj[c])
					if (b_move[o] == 0 && !vis[o]) st.push_back(o);
			}
		}
		pr[i] = b_move[i] == 0 && !asleep[i];
	}
	up(sl_asleep, asleep); up(sl_counter, counter); up(sl_ticks, ticks); up(sl_hlen, hlen); up(sl_hhead, hhead); up(sl_proc, pr);
	proc = true;
}

IslTables cp_world::isl_tables() {
	return IslTables{isl_root.p, isl_off.p, isl_edge_off.p, isl_body_off.p, isl_body_cnt.p, isl_stack_off.p, isl_joint_off.p};
}

// The last island schedule as (index, kind) rows and island offsets, read
// back from the device.
void cp_world::fetch_islands() {
	if (!islands_on_device) return;
	std::vector<int> offs, rows;
	down(offs, isl_off.p, (size_t)isl_n + 1);
	down(rows, isl_rows.p, 2 * (size_t)offs[isl_n]);
	last_rows.assign(rows.begin(), rows.end());
	last_offs.assign(offs.begin(), offs.end());
	islands_on_device = false;
}

// The colour schedule as one sequential row list per scene (colours in order,
// edges of a colour in order, contacts of an edge in cache order, normal row
// then friction row): edges of one colour share no dynamic body, so the
// parallel sweep equals this sequential one.
void cp_world::record_colored_schedule() {
	std::vector<int> order, co, nc, er, ef, ecnt, live;
	down(order, e_order.p, nedges);
	down(co, color_off.p, (size_t)nscenes * (CP_MAXC + 2));
	down(nc, ncolors.p, nscenes);
	down(er, edge_range.p, 2 * (size_t)nscenes);
	down(ef, e_first.p, nedges);
	down(ecnt, e_cnt.p, nedges);
	down(live, cb[cur].c.live, ncontacts);
	islands_on_device = false;
	last_rows.clear();
	last_offs.assign(1, 0);
	// first cache index of each scene's contacts (keys are lo << 32 | hi, sorted)
	std::vector<uint64_t> keys;
	down(keys, cb[cur].c.key, ncontacts);
	last_scene_coff.assign(nscenes + 1, ncontacts);
	for (int k = ncontacts - 1; k >= 0; k--) {
		int sc = b_scene[(size_t)(keys[k] >> 32)];
		for (int t = sc; t >= 0 && last_scene_coff[t] > k; t--) last_scene_coff[t] = k;
	}
	for (int s = 0; s < nscenes; s++) {
		for (int c = 0; c < nc[s]; c++) {
			const int* o = &co[(size_t)s * (CP_MAXC + 2)];
			for (int x = o[c]; x < o[c + 1]; x++) {
				int e = order[er[2 * s] + x];
				for (int k = ef[e]; k < ef[e] + ecnt[e]; k++) {
					if (!live[k]) continue;
					last_rows.push_back(k); last_rows.push_back(0);
					last_rows.push_back(k); last_rows.push_back(1);
				}
			}
		}
		last_offs.push_back((int64_t)last_rows.size() / 2);
	}
}

// Scenes too large for LDS: the global-memory kernel over those scenes only.
void cp_world::launch_pgs_colored_lds_global(const Bodies& b, const Params& P) {
	launch_pgs_colored(b, cb[cur].c, scene_body_off.p, nscenes, edge_range.p, e_desc.p, e_lmask.p, color_off.p, ncolors.p, P, cfg.solver_threads,
	                   stream, cls_dev[3].p, (int)cls_lists[3].size());
}

// One coloured PGS solve: the largest size class (and the global-memory one)
// on the world's stream, the two smaller classes beside it on the side
// streams.
void cp_world::colored_pgs(const Bodies& b, const Params& P) {
	const int* thr = cfg.pgs_class_threads;
	bool fork = !cls_lists[0].empty() || !cls_lists[1].empty();
	if (fork) {
		HIPCHK(hipEventRecord(fork_ev, stream));
		for (auto& x : side) HIPCHK(hipStreamWaitEvent(x, fork_ev, 0));
	}
	launch_pgs_colored_lds(b, cb[cur].c, cls_dev[2].p, (int)cls_lists[2].size(), cls_max[2], scene_body_off.p, edge_range.p, e_desc.p,
	                       e_lmask.p, color_off.p, ncolors.p, P, thr[2], stream);
	if (!cls_lists[3].empty()) launch_pgs_colored_lds_global(b, P);
	for (int c = 0; c < 2 && fork; c++)
		launch_pgs_colored_lds(b, cb[cur].c, cls_dev[c].p, (int)cls_lists[c].size(), cls_max[c], scene_body_off.p, edge_range.p, e_desc.p,
		                       e_lmask.p, color_off.p, ncolors.p, P, thr[c], side[c]);
	if (fork)
		for (int c = 0; c < 2; c++) {
			HIPCHK(hipEventRecord(join_ev[c], side[c]));
			HIPCHK(hipStreamWaitEvent(stream, join_ev[c], 0));
		}
}

// The edge colouring as one launch per PGS size class, each with LDS for its
// largest scene (one launch sized for the largest scene of all would fit a
// few small scenes' workgroups on a CU), the two smaller classes beside the
// others on the side streams.
void cp_world::color_by_class(const Bodies& b) {
	HIPCHK(hipEventRecord(fork_ev, stream));
	for (auto& x : side) HIPCHK(hipStreamWaitEvent(x, fork_ev, 0));
	hipStream_t st[4] = {side[0], side[1], stream, stream};
	for (int c = 3; c >= 0; c--)
		launch_color(b, scene_body_off.p, (int)cls_lists[c].size(), e_key.p, e_live.p, nedges, e_color.p, e_order.p, color_off.p, ncolors.p,
		             edge_range.p, cls_max[c], st[c], cls_dev[c].p);
	for (int c = 0; c < 2; c++) {
		HIPCHK(hipEventRecord(join_ev[c], side[c]));
		HIPCHK(hipStreamWaitEvent(stream, join_ev[c], 0));
	}
}

// The substep loop as one launch per size class (the largest class and the
// global-memory one on the world's stream, the two smaller classes beside it).
void cp_world::fused_substeps(const Bodies& b, const Params& P) {
	const int* thr = cfg.pgs_class_threads;
	bool fork = !cls_lists[0].empty() || !cls_lists[1].empty();
	if (fork) {
		HIPCHK(hipEventRecord(fork_ev, stream));
		for (auto& x : side) HIPCHK(hipStreamWaitEvent(x, fork_ev, 0));
	}
	const Contacts& C = cb[cur].c;
	launch_substeps_lds(b, C, cls_dev[2].p, (int)cls_lists[2].size(), cls_max[2], scene_body_off.p, edge_range.p, e_first.p, e_cnt.p,
	                    e_desc.p, e_lmask.p, color_off.p, ncolors.p, P, thr[2], stream);
	launch_substeps_global(b, C, cls_dev[3].p, (int)cls_lists[3].size(), scene_body_off.p, edge_range.p, e_first.p, e_cnt.p, e_desc.p,
	                       e_lmask.p, color_off.p, ncolors.p, P, cfg.solver_threads, stream);
	for (int c = 0; c < 2 && fork; c++)
		launch_substeps_lds(b, C, cls_dev[c].p, (int)cls_lists[c].size(), cls_max[c], scene_body_off.p, edge_range.p, e_first.p, e_cnt.p,
		                    e_desc.p, e_lmask.p, color_off.p, ncolors.p, P, thr[c], side[c]);
	if (fork)
		for (int c = 0; c < 2; c++) {
			HIPCHK(hipEventRecord(join_ev[c], side[c]));
			HIPCHK(hipStreamWaitEvent(stream, join_ev[c], 0));
		}
}

// Substep s of the coloured mode with the contact count read on the device.
void cp_world::colored_substep(const Bodies& b, const Params& P, int s, const int* ncp) {
	const int cap = cb[cur].c.cap;
	if (P.precision == CP_PREC_SPLIT && s == 0) launch_iinv32(b, stream);  // later substeps: k_integrate refreshed it
	launch_build_rows(b, cb[cur].c, cap, P, stream, ncp);
	colored_pgs(b, P);
	launch_rows_writeback(cb[cur].c, cap, P, stream, ncp);
	if (s == cfg.substeps - 1) launch_age(cb[cur].c, cap, stream, ncp);
	launch_integrate(b, P, stream);
}

// The substep loop as a HIP graph: captured the first time (and whenever a
// buffer, size or setting it depends on changes), then replayed; the step's
// contact count reaches the kernels through d_nc.
void cp_world::graph_substeps(const Bodies& b, const Params& P) {
	d_nc.reserve(1);
	launch_set_count(d_nc.p, ncontacts, stream);
	const Contacts& C = cb[cur].c;
	std::vector<uint64_t> key = {(uint64_t)(uintptr_t)C.key, (uint64_t)(uintptr_t)C.row, (uint64_t)C.cap, (uint64_t)(uintptr_t)C.live,
	                             (uint64_t)b.n, (uint64_t)b.nl, (uint64_t)(uintptr_t)b.com, (uint64_t)(uintptr_t)b.dv, (uint64_t)(uintptr_t)b.iinv32,
	                             (uint64_t)(uintptr_t)e_desc.p, (uint64_t)(uintptr_t)e_lmask.p, (uint64_t)(uintptr_t)color_off.p,
	                             (uint64_t)(uintptr_t)ncolors.p, (uint64_t)(uintptr_t)edge_range.p, (uint64_t)(uintptr_t)scene_body_off.p,
	                             (uint64_t)nscenes, (uint64_t)cfg.substeps, (uint64_t)cfg.solver_threads, (uint64_t)(uintptr_t)d_nc.p};
	for (int c = 0; c < 4; c++) {
		key.push_back((uint64_t)(uintptr_t)cls_dev[c].p);
		key.push_back((uint64_t)cls_lists[c].size());
		key.push_back((uint64_t)cls_max[c]);
	}
	for (int c = 0; c < 3; c++) key.push_back((uint64_t)cfg.pgs_class_threads[c]);
	{
		uint64_t pw[(sizeof(Params) + 7) / 8] = {};
		memcpy(pw, &P, sizeof(Params));
		key.insert(key.end(), pw, pw + (sizeof(Params) + 7) / 8);
	}
	Captured& g = graphs[cur];
	if (!g.exec || g.key != key) {
		if (g.exec) {
			HIPCHK(hipGraphExecDestroy(g.exec));
			g.exec = nullptr;
		}
		hipGraph_t gr = nullptr;
		HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeRelaxed));
		try {
			for (int s = 0; s < cfg.substeps; s++) colored_substep(b, P, s, d_nc.p);
		} catch (...) {
			(void)hipStreamEndCapture(stream, &gr);
			if (gr) (void)hipGraphDestroy(gr);
			throw;
		}
		HIPCHK(hipStreamEndCapture(stream, &gr));
		hipError_t e = hipGraphInstantiate(&g.exec, gr, nullptr, nullptr, 0);
		(void)hipGraphDestroy(gr);
		if (e != hipSuccess) {
			g.exec = nullptr;
			throw ApiErr{CP_ERR_DEVICE};
		}
		g.key = key;
	}
	HIPCHK(hipGraphLaunch(g.exec, stream));
}

void cp_world::step() {
	if (serving) serving_commit();
	if (!j_type.empty() && cfg.pgs_mode == CP_PGS_COLORED) throw ApiErr{CP_ERR_STATE};  // joints: island solves only
	if (cfg.sleeping && (cfg.pgs_mode == CP_PGS_COLORED || serving)) throw ApiErr{CP_ERR_STATE};  // sleeping: island solves only
	if (cfg.sleeping && std::count(j_kind.begin(), j_kind.end(), (int)CP_JOINT_SPRING)) throw ApiErr{CP_ERR_STATE};  // springs: no sleeping
	if (sl_active && !cfg.sleeping && cfg.pgs_mode == CP_PGS_COLORED) {
		// the coloured mode keeps no asleep bodies: all wake
		HIPCHK(hipMemsetAsync(sl_asleep.p, 0, (size_t)nb * sizeof(int), stream));
		HIPCHK(hipMemsetAsync(sl_asleep_rm.p, 0, (size_t)nb * sizeof(int), stream));
		sl_active = false;
	}
	if (cfg.sleeping) sl_active = true;
	bool sl_proc_set = false;
	if (sl_active) prepare_sleep(sl_proc_set);
	if (!hs_explicit && cfg.hol_iters > 0 && (hs_scan || hs_changed)) {
		discover_systems();
		hs_scan = false;
	}
	if (joints_dirty) upload_joints();
	if (drive_dirty) upload_drives();
	Params P = params();
	Bodies b = bodies();
	init_new_joints(b);
	schedule_springs();
	if (timing) HIPCHK(hipEventRecord(ev[0], stream));
	if (sl_active) {
		SleepDev SD = sleep_dev();
		if (sl_proc_set) SD.proc = sl_proc.p;
		launch_gravity_sleep(b, P, SD, stream);
	} else launch_gravity(b, P, stream);
	int npairs = broadphase();
	if (timing) HIPCHK(hipEventRecord(ev[1], stream));
	npairs = narrowphase(npairs);
	if (timing) HIPCHK(hipEventRecord(ev[2], stream));
	persist(npairs);
	edges();
	bool ordered = cfg.pgs_mode != CP_PGS_COLORED;  // island solves (reference order or coloured islands)
	Joints J = joints();
	HSysDev H = hsys();
	const bool hol = cfg.hol_iters > 0 && H.nsys > 0;
	// holonomic systems created or loaded from a state: rows from the current
	// state with the stored impulses, then the factor (PhysicsEngine.cpp:1507-1522)
	if (hol && hs_pending) {
		launch_joint_rows(b, J, P, P.sub_dt, 2, 0, stream);
		launch_hsys_factor(J, H, cfg.cfm, stream);
		hs_pending = false;
	}
	int nisl = 0;
	if (ordered) {
		schedule_islands(cfg.pgs_mode == CP_PGS_ISLAND_COLORED);
		nisl = isl_n;
		// the asleep flags when the reference removes expired contacts (the
		// first substep, after the island decision): the next step's merge
		// spares the contacts of these bodies
		if (sl_active) HIPCHK(hipMemcpyAsync(sl_asleep_rm.p, sl_asleep.p, (size_t)nb * sizeof(int), hipMemcpyDeviceToDevice, stream));
		if (hol) lin.reserve(6 * (size_t)lin_nl);
	} else {
		color_off.reserve((size_t)nscenes * (CP_MAXC + 2));
		ncolors.reserve(nscenes);
		edge_range.reserve(2 * (size_t)nscenes);
		e_desc.reserve(nedges);
		e_lmask.reserve(nedges);
		if (nedges) HIPCHK(hipMemsetAsync(e_order.p, 0xFF, (size_t)nedges * sizeof(int), stream));  // unused slots read -1
		if (serving || cls_lists.size() != 4)
			launch_color(b, scene_body_off.p, live_scenes, e_key.p, e_live.p, nedges, e_color.p, e_order.p, color_off.p, ncolors.p,
			             edge_range.p, max_scene_bodies, stream);
		else
			color_by_class(b);
		launch_edge_slots(b, cb[cur].c, e_order.p, nedges, e_first.p, e_cnt.p, e_desc.p, e_lmask.p, stream);
		if (cfg.diagnostics & 2) record_colored_schedule();
	}
	if (timing) HIPCHK(hipEventRecord(ev[3], stream));
	int npgs = 0;
	const bool use_graph = cfg.graph && !ordered && !(cfg.diagnostics & 1) && ncontacts > 0;
	const bool use_fused = cfg.fused && !cfg.graph && !ordered && !(cfg.diagnostics & 1);
	if (use_graph) graph_substeps(b, P);
	fused_step = use_fused;
	if (use_fused) {
		fused_substeps(b, P);
		launch_age(cb[cur].c, ncontacts, stream);
	}
	for (int s = 0; s < cfg.substeps && !use_graph && !use_fused; s++) {
		// spring impulses come before any row is built (PhysicsEngine.cpp:1313-1355)
		launch_springs(b, J, sp_ngroups, sp_goff.p, sp_list.p, P.sub_dt, stream);
		if (P.precision == CP_PREC_SPLIT && s == 0) launch_iinv32(b, stream);  // later substeps: k_integrate refreshed it
		launch_build_rows(b, cb[cur].c, ncontacts, P, stream);
		// joints; holonomic ones are current at the first substep (PhysicsEngine.cpp:1396-1403)
		launch_joint_rows(b, J, P, P.sub_dt, 0, s == 0, stream);
		bool tev = timing && npgs < 64;
		if (tev) HIPCHK(hipEventRecord(pgs_ev[2 * npgs], stream));
		if (ordered) {
			HIPCHK(hipMemsetAsync(dv.p, 0, 12 * (size_t)nb * sizeof(double), stream));
			HIPCHK(hipMemsetAsync(applied.p, 0, (size_t)nb * sizeof(int), stream));
			if (hol && s == 0) HIPCHK(hipMemsetAsync(lin.p, 0, 6 * (size_t)lin_nl * sizeof(double), stream));
			const bool icol = cfg.pgs_mode == CP_PGS_ISLAND_COLORED;
			IslandSched S{nisl, lin_nl, isl_rows.p, isl_off.p, isl_sys.p, isl_joint_off.p, isl_sys_cnt.p, isl_jlin.p,
			              isl_eoff.p, isl_coff.p, isl_icoff.p, isl_ncol.p, isl_serial.p,
			              icol ? isl_bod.p : nullptr, icol ? isl_body_off.p : nullptr, icol ? isl_nbod.p : nullptr,
			              icol ? isl_cloc.p : nullptr, icol ? isl_jloc.p : nullptr, icol ? isl_lds_bodies : 0,
			              ib_vlist.p, isl_body_off.p, isl_body_cnt.p, sleep_dev()};
			launch_pgs_islands(b, cb[cur].c, J, H, S, P, applied.p, lin.p, s == 0, cfg.pgs_mode == CP_PGS_ISLAND_COLORED, stream);
		} else
			colored_pgs(b, P);
		if (tev) {
			HIPCHK(hipEventRecord(pgs_ev[2 * npgs + 1], stream));
			npgs++;
		}
		launch_rows_writeback(cb[cur].c, ncontacts, P, stream);
		if (s == cfg.substeps - 1 && (cfg.diagnostics & 1)) {
			resid_acc.reserve(4);
			HIPCHK(hipMemsetAsync(resid_acc.p, 0, 4 * sizeof(double), stream));
			launch_pgs_residual(b, cb[cur].c, ncontacts, P, resid_acc.p, stream);
		}
		if (s == cfg.substeps - 1) launch_age(cb[cur].c, ncontacts, stream);
		launch_integrate(b, P, stream);
	}
	if (timing) HIPCHK(hipEventRecord(ev[4], stream));
	launch_refresh_geometry(b, shapes(), 1, stream);
	// calculateHolonomicSystemInversesAsync (PhysicsEngine.cpp:1615-1648): the
	// rows and factors the next step's first substep uses
	if (hol) {
		launch_joint_rows(b, J, P, P.sub_dt, 1, 0, stream);
		launch_hsys_factor(J, H, cfg.cfm, stream);
	}
	launch_check_finite(b, nf_body.p, nf_any.p, stream);
	if (cfg.diagnostics & 1) {
		double h[4];
		HIPCHK(hipMemcpy(h, resid_acc.p, 3 * sizeof(double), hipMemcpyDeviceToHost));
		last_resid[0] = std::sqrt(h[0]);
		last_resid[1] = std::sqrt(h[1]);
		last_resid[2] = h[1] > 0 ? std::sqrt(h[0] / h[1]) : 0.0;
		last_resid[3] = h[2];
	}
	if (timing) {
		HIPCHK(hipEventRecord(ev[5], stream));
		HIPCHK(hipEventSynchronize(ev[5]));
		for (int i = 0; i < 5; i++) {
			float ms = 0;
			HIPCHK(hipEventElapsedTime(&ms, ev[i], ev[i + 1]));
			last_ms[i] = ms;
		}
		double pg = 0;
		for (int i = 0; i < npgs; i++) {
			float ms = 0;
			HIPCHK(hipEventElapsedTime(&ms, pgs_ev[2 * i], pgs_ev[2 * i + 1]));
			pg += ms;
		}
		last_ms[5] = pg;
		last_ms[6] = npgs;
		if (fused_step) {  // the solve is inside the fused substep kernels: report their span
			last_ms[5] = last_ms[3];
			last_ms[6] = 0;
		}
		last_ms[7] = live_contacts;
	}
}

// ------------------------------------------------- snapshots and trajectories
void cp_world::record_frame() {
	const size_t n = (size_t)nb;
	if (traj_cap < 7 * n) {
		if (traj_host) HIPCHK(hipHostFree(traj_host));
		traj_host = nullptr;
		traj_cap = 0;
		HIPCHK(hipHostMalloc((void**)&traj_host, 7 * n * sizeof(double), hipHostMallocDefault));
		traj_cap = 7 * n;
	}
	HIPCHK(hipMemcpyAsync(traj_host, com.p, 3 * n * sizeof(double), hipMemcpyDeviceToHost, stream));
	HIPCHK(hipMemcpyAsync(traj_host + 3 * n, q.p, 4 * n * sizeof(double), hipMemcpyDeviceToHost, stream));
	HIPCHK(hipStreamSynchronize(stream));
	traj_frame.resize(7 * n);
	for (size_t i = 0; i < n; i++) {
		for (int c = 0; c < 3; c++) traj_frame[7 * i + c] = traj_host[c * n + i];
		for (int c = 0; c < 4; c++) traj_frame[7 * i + 3 + c] = traj_host[(3 + c) * n + i];
	}
	if (fwrite(traj_frame.data(), sizeof(double), 7 * n, traj) != 7 * n) throw ApiErr{CP_ERR_IO};
}

// Resumed admissions: the snapshot's world inertia over the freshly
// scattered bodies, then the snapshot's contacts merged into the cache (the
// rows the reference would still hold, ordered by body pair; the new scene's
// pairs are new to the cache, and rows keep their order within a pair).
void cp_world::apply_resumes() {
	if (resumes.empty()) return;
	HIPCHK(hipStreamSynchronize(stream));
	const size_t N = (size_t)nb;
	for (const Resume& r : resumes) {
		std::vector<double> comp((size_t)r.n);
		for (int c = 0; c < 18; c++) {
			for (int i = 0; i < r.n; i++) comp[i] = r.inertia[18 * (size_t)i + c];
			double* dst = (c < 9 ? I.p + c * N : Iinv.p + (c - 9) * N) + r.start;
			HIPCHK(hipMemcpy(dst, comp.data(), (size_t)r.n * sizeof(double), hipMemcpyHostToDevice));
		}
	}
	int nc = 0;
	int st = cp_world_get_contacts(this, 0, nullptr, nullptr, nullptr, &nc);
	if (st != CP_OK && st != CP_ERR_CAPACITY) throw ApiErr{st};
	std::vector<int64_t> ci(6 * (size_t)nc);
	std::vector<uint64_t> cm(2 * (size_t)nc);
	std::vector<double> cf(21 * (size_t)nc);
	if (nc && cp_world_get_contacts(this, nc, ci.data(), cm.data(), cf.data(), &nc) != CP_OK) throw ApiErr{CP_ERR_DEVICE};
	for (const Resume& r : resumes) {
		ci.insert(ci.end(), r.ci.begin(), r.ci.end());
		cm.insert(cm.end(), r.cm.begin(), r.cm.end());
		cf.insert(cf.end(), r.cf.begin(), r.cf.end());
	}
	const int tot = (int)(ci.size() / 6);
	std::vector<int> ord(tot);
	for (int k = 0; k < tot; k++) ord[k] = k;
	std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) {
		return ci[6 * (size_t)a] != ci[6 * (size_t)b] ? ci[6 * (size_t)a] < ci[6 * (size_t)b] : ci[6 * (size_t)a + 1] < ci[6 * (size_t)b + 1];
	});
	std::vector<int64_t> ci2(ci.size());
	std::vector<uint64_t> cm2(cm.size());
	std::vector<double> cf2(cf.size());
	for (int k = 0; k < tot; k++) {
		std::copy_n(&ci[6 * (size_t)ord[k]], 6, &ci2[6 * (size_t)k]);
		std::copy_n(&cm[2 * (size_t)ord[k]], 2, &cm2[2 * (size_t)k]);
		std::copy_n(&cf[21 * (size_t)ord[k]], 21, &cf2[21 * (size_t)k]);
	}
	resumes.clear();
	st = cp_world_set_contacts(this, tot, ci2.data(), cm2.data(), cf2.data());
	if (st != CP_OK) throw ApiErr{st};
}

namespace {
const char SNAP_MAGIC[8] = {'C', 'P', 'S', 'N', 'A', 'P', '2', '\n'};
uint64_t fnv1a(uint64_t h, const void* p, size_t n) {
	const unsigned char* b = static_cast<const unsigned char*>(p);
	for (size_t i = 0; i < n; i++) {
		h ^= b[i];
		h *= 1099511628211ull;
	}
	return h;
}
struct SnapFile {  // a snapshot as written: counts, then the sections in file order
	int64_t counts[4] = {0, 0, 0, 0};  // bodies, contacts, joints, sleep (0 none, else history slots + 1)
	std::vector<double> state, inertia, cf, joint, jdrive, jrow, jpimp, sl_counter, sl_hist;
	std::vector<int64_t> ci, jproj, sl_ints;
	std::vector<uint64_t> cm;
};
}  // namespace

// ------------------------------------------------- narrowphase without allocation
// cp_narrowphase_noalloc runs the narrowphase kernels on a fixed device array
// instead of the world's grow-only buffers: pair keys, manifold slots (stride
// NA_MAX_PAIRS), per-class pair lists and counts, and the exported heads.
// Every call on a device shares that device's copy; it keeps the last call's
// contents (the class counts are cleared by the kernels, every slot of the
// call's pairs is written), so calls that overlap on one device race.
#define NA_MAX_PAIRS 16384
namespace {
constexpr size_t NA_PAIRS = 0;                               // uint64 [MAX]
constexpr size_t NA_COUNT = NA_PAIRS + 8 * NA_MAX_PAIRS;     // int [MAX]
constexpr size_t NA_NORMAL = NA_COUNT + 4 * NA_MAX_PAIRS;    // double [3][MAX]
constexpr size_t NA_MAXPEN = NA_NORMAL + 24 * NA_MAX_PAIRS;  // double [MAX]
constexpr size_t NA_PTS = NA_MAXPEN + 8 * NA_MAX_PAIRS;      // double [MAX][4][4]
constexpr size_t NA_MAGIC = NA_PTS + 128 * NA_MAX_PAIRS;     // uint64 [MAX][4][2]
constexpr size_t NA_MAT = NA_MAGIC + 64 * NA_MAX_PAIRS;      // double [MAX * 4][3]
constexpr size_t NA_HEADS = NA_MAT + 96 * NA_MAX_PAIRS;      // double [MAX][4] (first the input pairs, int64 [MAX][2])
constexpr size_t NA_LISTS = NA_HEADS + 32 * NA_MAX_PAIRS;    // int [7][MAX]
constexpr size_t NA_CLASSCNT = NA_LISTS + 28 * NA_MAX_PAIRS; // int [8]
constexpr size_t NA_BYTES = NA_CLASSCNT + 256;
constexpr int NA_MAX_DEVICES = 64;
void* na_ptr[NA_MAX_DEVICES] = {};  // per device: the address of its copy of `scratch`
std::mutex na_mu;
}  // namespace
static __device__ __align__(256) unsigned char scratch[NA_BYTES];

// Once per device, with that device current: the device address of its copy
// of `scratch` (the host-side name is only a stand-in for it).
static void na_setup(int dev) {
	if (dev < 0 || dev >= NA_MAX_DEVICES) return;
	std::lock_guard<std::mutex> lk(na_mu);
	if (na_ptr[dev]) return;
	void* p = nullptr;
	HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
	na_ptr[dev] = p;
}

__global__ void k_na_keys(const int64_t* ids, int n, uint64_t* keys) {
	int k = blockIdx.x * blockDim.x + threadIdx.x;
	if (k < n) keys[k] = ((uint64_t)(ids[2 * k] - 1) << 32) | (uint64_t)(ids[2 * k + 1] - 1);
}
__global__ void k_na_heads(const int* count, const double* normal, const double* maxpen, int cap, int n, double* heads) {
	int k = blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= n) return;
	for (int c = 0; c < 3; c++) heads[4 * k + c] = count[k] ? normal[(size_t)c * cap + k] : 0.0;
	heads[4 * k + 3] = maxpen[k];
}

// ============================================================================ C ABI
static int guard(cp_world* w, const std::function<void()>& f) {
	try {
		if (w) HIPCHK(hipSetDevice(w->device));
		f();
		return CP_OK;
	} catch (const ApiErr& e) {
		return e.code;
	} catch (const DevErr& e) {
		fprintf(stderr, "cp_hip: %s failed: %s\n", e.what, hipGetErrorString(e.e));
		return CP_ERR_DEVICE;
	} catch (const std::bad_alloc&) {
		return CP_ERR_OOM;
	}
}

extern "C" {

void cp_config_default(cp_config* c) {
	c->dt = 1.0 / 90;
	c->gravity[0] = 0; c->gravity[1] = -6; c->gravity[2] = 0;
	c->vel_iters = 4; c->pos_iters = 1; c->hol_iters = 1;
	c->substeps = 8;
	c->cfm = 0.0000025;
	c->contact_hardness = 1000;
	c->contact_life = 6;
	c->gyro_ticks = 4;
	c->pgs_mode = CP_PGS_COLORED;
	c->solver_threads = 1024;
	c->diagnostics = 0;
	c->defect_fixes = 3;
	c->pgs_class_bodies[0] = 192;
	c->pgs_class_bodies[1] = 768;
	c->pgs_class_bodies[2] = 3072;
	c->pgs_class_threads[0] = 64;
	c->pgs_class_threads[1] = 128;
	c->pgs_class_threads[2] = 256;
	c->graph = 0;
	c->fused = 0;
	c->precision = 0;
	c->sleeping = 0;
	c->sleep_tick_threshold = 3;
	c->sleep_vel_coeff = 0.1;
	c->sleep_ang_vel_eps = 0.1;
	c->sleep_accel_coeff = 0.044;
	c->sleep_delay = 0.5;
}

const char* cp_status_string(int s) {
	switch (s) {
	case CP_OK: return "ok";
	case CP_ERR_INVALID: return "invalid argument";
	case CP_ERR_OOM: return "device out of memory";
	case CP_ERR_DEVICE: return "HIP runtime error";
	case CP_ERR_CAPACITY: return "buffer too small";
	case CP_ERR_STATE: return "invalid world state";
	case CP_ERR_NONFINITE: return "non-finite body state (NaN or infinity)";
	case CP_ERR_IO: return "file could not be read or written, or is not a valid file of its kind";
	}
	return "unknown";
}

static bool valid_threads(int t) { return t >= 64 && t <= 1024 && (t % 64) == 0; }
static bool valid_cfg(const cp_config* c) {
	if (!c) return false;
	const int* lim = c->pgs_class_bodies;
	bool classes = lim[0] >= 1 && lim[0] < lim[1] && lim[1] < lim[2] && lim[2] <= 3072;
	for (int k = 0; k < 3; k++) classes = classes && valid_threads(c->pgs_class_threads[k]);
	return c->dt > 0 && c->substeps >= 1 && c->vel_iters >= 0 && c->pos_iters >= 0 && c->gyro_ticks >= 1 && valid_threads(c->solver_threads) &&
	       (c->pgs_mode >= 0 && c->pgs_mode <= 2) && (c->precision == CP_PREC_FP64 || c->precision == CP_PREC_SPLIT) && classes &&
	       (c->graph == 0 || c->graph == 1) && (c->fused == 0 || c->fused == 1) && (c->sleeping == 0 || c->sleeping == 1) &&
	       c->sleep_tick_threshold >= 1 && c->sleep_delay >= 0 && c->sleep_delay / c->dt < 1e5 && std::isfinite(c->sleep_vel_coeff) &&
	       std::isfinite(c->sleep_ang_vel_eps) && std::isfinite(c->sleep_accel_coeff);
}

int cp_world_create(const cp_config* cfg, cp_world** out) {
	if (!out || !valid_cfg(cfg)) return CP_ERR_INVALID;
	*out = nullptr;
	int ndev = 0;
	if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return CP_ERR_DEVICE;
	cp_world* w = new cp_world();
	w->cfg = *cfg;
	int st = guard(w, [&] {
		HIPCHK(hipGetDevice(&w->device));
		HIPCHK(hipStreamCreate(&w->stream));
		for (auto& x : w->side) HIPCHK(hipStreamCreate(&x));
		HIPCHK(hipEventCreateWithFlags(&w->fork_ev, hipEventDisableTiming));
		for (auto& e : w->join_ev) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
		w->fk = Fork{{w->side[0], w->side[1]}, w->fork_ev, {w->join_ev[0], w->join_ev[1]}};
		for (auto& e : w->ev) HIPCHK(hipEventCreate(&e));
		for (auto& e : w->pgs_ev) HIPCHK(hipEventCreate(&e));
		na_setup(w->device);
	});
	if (st != CP_OK) {
		delete w;
		return st;
	}
	*out = w;
	return CP_OK;
}

int cp_world_destroy(cp_world* w) {
	if (!w) return CP_ERR_INVALID;
	(void)hipSetDevice(w->device);
	(void)hipStreamSynchronize(w->stream);
	w->free_all();
	for (auto& e : w->ev) (void)hipEventDestroy(e);
	for (auto& e : w->pgs_ev) (void)hipEventDestroy(e);
	for (auto& g : w->graphs)
		if (g.exec) (void)hipGraphExecDestroy(g.exec);
	for (auto& x : w->side)
		if (x) {
			(void)hipStreamSynchronize(x);
			(void)hipStreamDestroy(x);
		}
	if (w->fork_ev) (void)hipEventDestroy(w->fork_ev);
	for (auto& e : w->join_ev)
		if (e) (void)hipEventDestroy(e);
	(void)hipStreamDestroy(w->stream);
	if (w->nf_host) (void)hipHostFree(w->nf_host);
	if (w->traj) fclose(w->traj);
	if (w->traj_host) (void)hipHostFree(w->traj_host);
	delete w;
	return CP_OK;
}

int cp_world_set_config(cp_world* w, const cp_config* cfg) {
	if (!w || !valid_cfg(cfg)) return CP_ERR_INVALID;
	w->cfg = *cfg;
	if (w->uploaded && !w->cls_sizes.empty() &&
	    !std::equal(w->cls_built, w->cls_built + 3, w->cfg.pgs_class_bodies))
		return guard(w, [&] { w->build_classes(); });
	return CP_OK;
}

int cp_shape_add_hull_exact(cp_world* w, int n_pts, const double* pts, int n_faces, const int64_t* face_off, const int64_t* face_verts,
                            const int64_t* surface_id, const int64_t* normal_calc, const int64_t* normal_dir, const double* gauss_normal,
                            const int64_t* gauss_red, const int64_t* gauss_ints, const double* gauss_vals, int n_arcs, const int64_t* arcs,
                            int n_edges, const int64_t* edges, int* shape_id) {
	if (!w || !shape_id || n_pts < 4 || n_faces < 4 || !pts || !face_off || !face_verts) return CP_ERR_INVALID;
	HullData d;
	d.pts.assign(pts, pts + 3 * (size_t)n_pts);
	for (int f = 0; f < n_faces; f++) {
		std::vector<int> fv;
		for (int64_t k = face_off[f]; k < face_off[f + 1]; k++) {
			if (face_verts[k] < 0 || face_verts[k] >= n_pts) return CP_ERR_INVALID;
			fv.push_back((int)face_verts[k]);
		}
		if (fv.size() < 3) return CP_ERR_INVALID;
		d.faces.push_back(fv);
		d.sid.push_back((int)surface_id[f]);
		d.ncalc.push_back((int)normal_calc[f]);
		d.ndir.push_back((int)normal_dir[f]);
		for (int c = 0; c < 3; c++) d.gnorm.push_back(gauss_normal[3 * f + c]);
		d.red.push_back((int)gauss_red[f]);
		d.minpid.push_back((int)gauss_ints[3 * f]);
		d.maxpid.push_back((int)gauss_ints[3 * f + 1]);
		d.refidx.push_back((int)gauss_ints[3 * f + 2]);
		d.gmin.push_back(gauss_vals[3 * f]);
		d.gmax.push_back(gauss_vals[3 * f + 1]);
		d.grefval.push_back(gauss_vals[3 * f + 2]);
	}
	for (int a = 0; a < n_arcs; a++) d.arcs.push_back({(int)arcs[2 * a], (int)arcs[2 * a + 1]});
	for (int e = 0; e < n_edges; e++) d.edges.push_back({(int)edges[2 * e], (int)edges[2 * e + 1]});
	*shape_id = w->hs.add_hull(d);
	if (!w->serving) w->uploaded = false;  // a serving world appends shape tables at its next commit
	return CP_OK;
}

int cp_shape_add_sphere(cp_world* w, const double* center, double radius, int* shape_id) {
	if (!w || !center || !shape_id || !(radius > 0)) return CP_ERR_INVALID;
	*shape_id = w->hs.add_sphere(center, radius);
	if (!w->serving) w->uploaded = false;  // a serving world appends shape tables at its next commit
	return CP_OK;
}

int cp_shape_add_cylinder(cp_world* w, const double* params8, const double* gauss42, const double* top24, const double* bot24,
                          int* shape_id) {
	if (!w || !params8 || !gauss42 || !top24 || !bot24 || !shape_id || !(params8[6] > 0) || !(params8[7] > 0)) return CP_ERR_INVALID;
	*shape_id = w->hs.add_cylinder(params8, gauss42, top24, bot24);
	if (!w->serving) w->uploaded = false;  // a serving world appends shape tables at its next commit
	return CP_OK;
}

int cp_cylinder_build(const double* center, double radius, double height, double* params8, double* gauss42, double* top24,
                      double* bot24) {
	if (!center || !params8 || !gauss42 || !top24 || !bot24 || !(radius > 0) || !(height > 0)) return CP_ERR_INVALID;
	build_cylinder(center, radius, height, params8, gauss42, top24, bot24);
	return CP_OK;
}

int cp_shape_add_mesh(cp_world* w, int n_pts, const double* pts, int n_tris, const int64_t* tri_idx, const double* tri_mat3, int* shape_id) {
	if (!w || n_pts < 3 || n_tris < 1 || !pts || !tri_idx || !tri_mat3 || !shape_id) return CP_ERR_INVALID;
	for (int k = 0; k < 3 * n_tris; k++)
		if (tri_idx[k] < 0 || tri_idx[k] >= n_pts) return CP_ERR_INVALID;
	for (int k = 0; k < 3 * n_pts; k++)
		if (!std::isfinite(pts[k])) return CP_ERR_INVALID;
	*shape_id = w->hs.add_mesh(build_mesh(n_pts, pts, n_tris, tri_idx, tri_mat3));
	if (!w->serving) w->uploaded = false;  // a serving world appends shape tables at its next commit
	return CP_OK;
}

int cp_shape_add_polyhedron(cp_world* w, int n_pts, const double* pts, int n_faces, const int64_t* face_off, const int64_t* face_verts,
                            int* shape_id) {
	if (!w || !shape_id || n_pts < 4 || n_faces < 4 || !pts || !face_off || !face_verts) return CP_ERR_INVALID;
	std::vector<std::vector<int>> faces;
	for (int f = 0; f < n_faces; f++) {
		std::vector<int> fv;
		for (int64_t k = face_off[f]; k < face_off[f + 1]; k++) {
			if (face_verts[k] < 0 || face_verts[k] >= n_pts) return CP_ERR_INVALID;
			fv.push_back((int)face_verts[k]);
		}
		if (fv.size() < 3) return CP_ERR_INVALID;
		faces.push_back(fv);
	}
	HullData d = build_polyhedron(std::vector<double>(pts, pts + 3 * (size_t)n_pts), faces);
	*shape_id = w->hs.add_hull(d);
	if (!w->serving) w->uploaded = false;  // a serving world appends shape tables at its next commit
	return CP_OK;
}

int cp_polyhedron_build(int n_pts, const double* pts, int n_faces, const int64_t* face_off, const int64_t* face_verts,
                        int64_t* fvo, int64_t* sid, int64_t* ncalc, int64_t* ndir, double* gn, int64_t* gred, int64_t* gint, double* gval,
                        int cap_pairs, int64_t* arcs_out, int* n_arcs, int64_t* edges_out, int* n_edges) {
	if (n_pts < 4 || n_faces < 4 || !pts || !face_off || !face_verts || !fvo || !sid || !ncalc || !ndir || !gn || !gred || !gint || !gval ||
	    !n_arcs || !n_edges)
		return CP_ERR_INVALID;
	std::vector<std::vector<int>> faces;
	for (int f = 0; f < n_faces; f++) {
		std::vector<int> fv;
		for (int64_t k = face_off[f]; k < face_off[f + 1]; k++) {
			if (face_verts[k] < 0 || face_verts[k] >= n_pts) return CP_ERR_INVALID;
			fv.push_back((int)face_verts[k]);
		}
		if (fv.size() < 3) return CP_ERR_INVALID;
		faces.push_back(fv);
	}
	HullData d = build_polyhedron(std::vector<double>(pts, pts + 3 * (size_t)n_pts), faces);
	size_t o = 0;
	for (int f = 0; f < n_faces; f++) {
		for (int v : d.faces[f]) fvo[o++] = v;
		sid[f] = d.sid[f]; ncalc[f] = d.ncalc[f]; ndir[f] = d.ndir[f]; gred[f] = d.red[f];
		for (int c = 0; c < 3; c++) gn[3 * f + c] = d.gnorm[3 * f + c];
		gint[3 * f] = d.minpid[f]; gint[3 * f + 1] = d.maxpid[f]; gint[3 * f + 2] = d.refidx[f];
		gval[3 * f] = d.gmin[f]; gval[3 * f + 1] = d.gmax[f]; gval[3 * f + 2] = d.grefval[f];
	}
	*n_arcs = (int)d.arcs.size();
	*n_edges = (int)d.edges.size();
	if ((int)d.arcs.size() > cap_pairs || (int)d.edges.size() > cap_pairs || !arcs_out || !edges_out) return CP_ERR_CAPACITY;
	for (size_t a = 0; a < d.arcs.size(); a++) { arcs_out[2 * a] = d.arcs[a].first; arcs_out[2 * a + 1] = d.arcs[a].second; }
	for (size_t e = 0; e < d.edges.size(); e++) { edges_out[2 * e] = d.edges[e].first; edges_out[2 * e + 1] = d.edges[e].second; }
	return CP_OK;
}

int cp_polyhedron_mass_properties(int n_pts, const double* pts, int n_faces, const int64_t* face_off, const int64_t* face_verts,
                                  double density, double* out13) {
	if (n_pts < 4 || n_faces < 4 || !pts || !face_off || !face_verts || !out13 || !(density > 0)) return CP_ERR_INVALID;
	std::vector<std::vector<int>> faces;
	for (int f = 0; f < n_faces; f++) {
		std::vector<int> fv;
		for (int64_t k = face_off[f]; k < face_off[f + 1]; k++) {
			if (face_verts[k] < 0 || face_verts[k] >= n_pts) return CP_ERR_INVALID;
			fv.push_back((int)face_verts[k]);
		}
		if (fv.size() < 3) return CP_ERR_INVALID;
		faces.push_back(fv);
	}
	polyhedron_mass_properties(std::vector<double>(pts, pts + 3 * (size_t)n_pts), faces, density, out13, out13 + 1, out13 + 10);
	return CP_OK;
}

int cp_polyhedron_body(int n_pts, const double* pts, int n_faces, const int64_t* face_off, const int64_t* face_verts, double density,
                       double* pts_out, double* mp19, double* com3) {
	if (n_pts < 4 || n_faces < 4 || !pts || !face_off || !face_verts || !pts_out || !mp19 || !com3 || !(density > 0))
		return CP_ERR_INVALID;
	std::vector<std::vector<int>> faces;
	for (int f = 0; f < n_faces; f++) {
		std::vector<int> fv;
		for (int64_t k = face_off[f]; k < face_off[f + 1]; k++) {
			if (face_verts[k] < 0 || face_verts[k] >= n_pts) return CP_ERR_INVALID;
			fv.push_back((int)face_verts[k]);
		}
		if (fv.size() < 3) return CP_ERR_INVALID;
		faces.push_back(fv);
	}
	std::vector<double> p(pts, pts + 3 * (size_t)n_pts);
	double t9[9];
	polyhedron_mass_properties(p, faces, density, mp19, t9, com3);
	HullData d = polyhedron_body(p, faces, density, mp19);
	for (size_t k = 0; k < d.pts.size(); k++) pts_out[k] = d.pts[k];
	return CP_OK;
}

int cp_body_add(cp_world* w, int shape_id, int movement, const double* mp, const double* material, int scene, int* body_index) {
	if (!w || !mp || !material || !body_index) return CP_ERR_INVALID;
	if (w->serving) return CP_ERR_STATE;  // serving worlds take whole scenes (cp_world_admit_scene)
	if (shape_id < 0 || shape_id >= (int)w->hs.type.size() || (movement != 0 && movement != 1)) return CP_ERR_INVALID;
	if (w->hs.type[shape_id] == SHAPE_MESH && movement != 1) return CP_ERR_INVALID;  // meshes are static (RigidBody.cpp:121)
	if (!w->b_scene.empty() && scene < w->b_scene.back()) return CP_ERR_INVALID;  // scenes are contiguous and ascending
	if (w->uploaded && w->ncontacts > 0) return CP_ERR_STATE;
	w->b_shape.push_back(shape_id);
	w->b_move.push_back(movement);
	w->b_scene.push_back(scene);
	w->b_massprops.insert(w->b_massprops.end(), mp, mp + 19);
	w->b_mat.insert(w->b_mat.end(), material, material + 4);
	*body_index = w->nb++;
	if (!w->serving) w->uploaded = false;  // a serving world appends shape tables at its next commit
	return CP_OK;
}

int cp_body_add_many(cp_world* w, int n, const int* shape_ids, const int* movement, const double* mp, const double* material,
                     const int* scene, int* first_index) {
	if (!w || n < 0 || !first_index || (n > 0 && (!shape_ids || !movement || !mp || !material || !scene))) return CP_ERR_INVALID;
	if (w->serving) return CP_ERR_STATE;
	if (w->uploaded && w->ncontacts > 0) return CP_ERR_STATE;
	int last = w->b_scene.empty() ? -1 : w->b_scene.back();
	for (int i = 0; i < n; i++) {
		if (shape_ids[i] < 0 || shape_ids[i] >= (int)w->hs.type.size() || (movement[i] != 0 && movement[i] != 1)) return CP_ERR_INVALID;
		if (w->hs.type[shape_ids[i]] == SHAPE_MESH && movement[i] != 1) return CP_ERR_INVALID;
		if (scene[i] < last) return CP_ERR_INVALID;
		last = scene[i];
	}
	*first_index = w->nb;
	w->b_shape.insert(w->b_shape.end(), shape_ids, shape_ids + n);
	w->b_move.insert(w->b_move.end(), movement, movement + n);
	w->b_scene.insert(w->b_scene.end(), scene, scene + n);
	w->b_massprops.insert(w->b_massprops.end(), mp, mp + 19 * (size_t)n);
	w->b_mat.insert(w->b_mat.end(), material, material + 4 * (size_t)n);
	w->nb += n;
	if (!w->serving) w->uploaded = false;  // a serving world appends shape tables at its next commit
	return CP_OK;
}

int cp_world_body_count(cp_world* w, int* n) {
	if (!w || !n) return CP_ERR_INVALID;
	*n = w->nb;
	return CP_OK;
}

int cp_world_set_state(cp_world* w, const double* s, const double* inertia) {
	if (!w || !s || w->nb == 0) return CP_ERR_INVALID;
	return guard(w, [&] {
		if (!w->uploaded) w->upload();
		int n = w->nb;
		std::vector<double> com(3 * (size_t)n), q(4 * (size_t)n), v(3 * (size_t)n), av(3 * (size_t)n);
		for (int i = 0; i < n; i++) {
			const double* r = s + 13 * (size_t)i;
			for (int c = 0; c < 3; c++) {
				com[c * (size_t)n + i] = r[c];
				v[c * (size_t)n + i] = r[7 + c];
				av[c * (size_t)n + i] = r[10 + c];
			}
			for (int c = 0; c < 4; c++) q[c * (size_t)n + i] = r[3 + c];
		}
		HIPCHK(hipMemcpy(w->com.p, com.data(), com.size() * 8, hipMemcpyHostToDevice));
		HIPCHK(hipMemcpy(w->q.p, q.data(), q.size() * 8, hipMemcpyHostToDevice));
		HIPCHK(hipMemcpy(w->vel.p, v.data(), v.size() * 8, hipMemcpyHostToDevice));
		HIPCHK(hipMemcpy(w->angv.p, av.data(), av.size() * 8, hipMemcpyHostToDevice));
		HIPCHK(hipMemcpy(w->prev_com.p, com.data(), com.size() * 8, hipMemcpyHostToDevice));
		HIPCHK(hipMemcpy(w->prev_q.p, q.data(), q.size() * 8, hipMemcpyHostToDevice));
		HIPCHK(hipMemset(w->pvel.p, 0, 3 * (size_t)n * 8));
		HIPCHK(hipMemset(w->pang.p, 0, 3 * (size_t)n * 8));
		if (inertia) {
			std::vector<double> I(9 * (size_t)n), Ii(9 * (size_t)n);
			for (int i = 0; i < n; i++)
				for (int c = 0; c < 9; c++) {
					I[c * (size_t)n + i] = inertia[18 * (size_t)i + c];
					Ii[c * (size_t)n + i] = inertia[18 * (size_t)i + 9 + c];
				}
			HIPCHK(hipMemcpy(w->I.p, I.data(), I.size() * 8, hipMemcpyHostToDevice));
			HIPCHK(hipMemcpy(w->Iinv.p, Ii.data(), Ii.size() * 8, hipMemcpyHostToDevice));
		} else {
			launch_update_inertia(w->bodies(), w->stream);
		}
		launch_refresh_geometry(w->bodies(), w->shapes(), 0, w->stream);
		launch_fill_int(w->nf_body.p, w->nb, 0, w->stream);  // a new state: no body has gone non-finite yet
		HIPCHK(hipStreamSynchronize(w->stream));
		w->place_meshes(s);
	});
}

int cp_world_get_state(cp_world* w, double* s) {
	if (!w || !s || !w->uploaded) return CP_ERR_INVALID;
	return guard(w, [&] {
		HIPCHK(hipStreamSynchronize(w->stream));
		int n = w->nb;
		std::vector<double> com, q, v, av;
		down(com, w->com.p, 3 * (size_t)n);
		down(q, w->q.p, 4 * (size_t)n);
		down(v, w->vel.p, 3 * (size_t)n);
		down(av, w->angv.p, 3 * (size_t)n);
		for (int i = 0; i < n; i++) {
			double* r = s + 13 * (size_t)i;
			for (int c = 0; c < 3; c++) {
				r[c] = com[c * (size_t)n + i];
				r[7 + c] = v[c * (size_t)n + i];
				r[10 + c] = av[c * (size_t)n + i];
			}
			for (int c = 0; c < 4; c++) r[3 + c] = q[c * (size_t)n + i];
		}
	});
}

int cp_world_get_inertia(cp_world* w, double* inertia) {
	if (!w || !inertia || !w->uploaded) return CP_ERR_INVALID;
	return guard(w, [&] {
		HIPCHK(hipStreamSynchronize(w->stream));
		int n = w->nb;
		std::vector<double> I, Ii;
		down(I, w->I.p, 9 * (size_t)n);
		down(Ii, w->Iinv.p, 9 * (size_t)n);
		for (int i = 0; i < n; i++)
			for (int c = 0; c < 9; c++) {
				inertia[18 * (size_t)i + c] = I[c * (size_t)n + i];
				inertia[18 * (size_t)i + 9 + c] = Ii[c * (size_t)n + i];
			}
	});
}

int cp_world_get_aabbs(cp_world* w, double* a) {
	if (!w || !a || !w->uploaded) return CP_ERR_INVALID;
	return guard(w, [&] {
		HIPCHK(hipStreamSynchronize(w->stream));
		int n = w->nb;
		std::vector<double> h;
		down(h, w->aabb.p, 6 * (size_t)n);
		for (int i = 0; i < n; i++)
			for (int c = 0; c < 6; c++) a[6 * (size_t)i + c] = h[c * (size_t)n + i];
	});
}

int cp_world_set_contacts(cp_world* w, int n, const int64_t* ii, const uint64_t* mg, const double* ff) {
	if (!w || n < 0 || (n > 0 && (!ii || !mg || !ff)) || !w->uploaded) return CP_ERR_INVALID;
	for (int k = 0; k < n; k++) {
		const int64_t* r = ii + 6 * (size_t)k;
		if (r[0] < 1 || r[1] <= r[0] || r[1] > w->nb) return CP_ERR_INVALID;
		if (k > 0) {
			const int64_t* p = ii + 6 * (size_t)(k - 1);
			if (p[0] > r[0] || (p[0] == r[0] && p[1] > r[1])) return CP_ERR_INVALID;  // grouped by sorted pair
		}
	}
	return guard(w, [&] {
		ContactBuf& B = w->cb[w->cur];
		B.reserve(std::max(n, 1));
		int K = B.cap;
		std::vector<uint64_t> key(n), cid(n), bid(n), seq(n);
		std::vector<double> lp1(3 * (size_t)K), lp2(3 * (size_t)K), ln1(3 * (size_t)K), fu(3 * (size_t)K), fw(3 * (size_t)K), imp_t(2 * (size_t)K);
		std::vector<double> pen(n), bounce(n), mu(n), imp_n(n);
		std::vector<int> life(n), live(n), sr(n);
		for (int k = 0; k < n; k++) {
			const int64_t* r = ii + 6 * (size_t)k;
			const double* f = ff + 21 * (size_t)k;
			key[k] = ((uint64_t)(r[0] - 1) << 32) | (uint64_t)(r[1] - 1);
			cid[k] = mg[2 * k];
			bid[k] = mg[2 * k + 1];
			seq[k] = (uint64_t)k;
			life[k] = (int)r[2];
			live[k] = (int)r[3];
			sr[k] = (int)r[4];
			for (int c = 0; c < 3; c++) {
				lp1[c * (size_t)K + k] = f[c];
				lp2[c * (size_t)K + k] = f[3 + c];
				ln1[c * (size_t)K + k] = f[6 + c];
				fu[c * (size_t)K + k] = f[15 + c];
				fw[c * (size_t)K + k] = f[18 + c];
			}
			pen[k] = f[9]; bounce[k] = f[10]; mu[k] = f[11]; imp_n[k] = f[12];
			imp_t[k] = f[13];
			imp_t[K + k] = f[14];
		}
		auto cp = [&](void* d, const void* h, size_t bytes) {
			if (bytes) HIPCHK(hipMemcpy(d, h, bytes, hipMemcpyHostToDevice));
		};
		Contacts& C = B.c;
		cp(C.key, key.data(), 8 * (size_t)n); cp(C.cid, cid.data(), 8 * (size_t)n); cp(C.bid, bid.data(), 8 * (size_t)n);
		cp(C.seq, seq.data(), 8 * (size_t)n);
		cp(C.lp1, lp1.data(), lp1.size() * 8); cp(C.lp2, lp2.data(), lp2.size() * 8); cp(C.ln1, ln1.data(), ln1.size() * 8);
		cp(C.fu, fu.data(), fu.size() * 8); cp(C.fw, fw.data(), fw.size() * 8); cp(C.imp_t, imp_t.data(), imp_t.size() * 8);
		cp(C.pen, pen.data(), 8 * (size_t)n); cp(C.bounce, bounce.data(), 8 * (size_t)n); cp(C.mu, mu.data(), 8 * (size_t)n);
		cp(C.imp_n, imp_n.data(), 8 * (size_t)n);
		cp(C.life, life.data(), 4 * (size_t)n); cp(C.live, live.data(), 4 * (size_t)n); cp(C.sready, sr.data(), 4 * (size_t)n);
		w->ncontacts = n;
	});
}

int cp_world_get_contact_count(cp_world* w, int* n) {
	if (!w || !n) return CP_ERR_INVALID;
	return guard(w, [&] {
		HIPCHK(hipStreamSynchronize(w->stream));
		std::vector<int> life;
		down(life, w->cb[w->cur].c.life, w->ncontacts);
		std::vector<char> gone = w->removed_contacts(life);
		int c = 0;
		for (char g2 : gone) c += !g2;
		*n = c;
	});
}

// Contacts that the reference would still hold (not removed at the last
// step's first substep), in cache order.
int cp_world_get_contacts(cp_world* w, int cap, int64_t* ii, uint64_t* mg, double* ff, int* nout) {
	if (!w || !nout) return CP_ERR_INVALID;
	int status = CP_OK;
	int g = guard(w, [&] {
		HIPCHK(hipStreamSynchronize(w->stream));
		const Contacts& C = w->cb[w->cur].c;
		int n = w->ncontacts, K = C.cap;
		std::vector<uint64_t> key, cid, bid;
		std::vector<double> lp1, lp2, ln1, fu, fw, imp_t, pen, bounce, mu, imp_n;
		std::vector<int> life, live, sr;
		down(key, C.key, n); down(cid, C.cid, n); down(bid, C.bid, n);
		down(lp1, C.lp1, 3 * (size_t)K); down(lp2, C.lp2, 3 * (size_t)K); down(ln1, C.ln1, 3 * (size_t)K);
		down(fu, C.fu, 3 * (size_t)K); down(fw, C.fw, 3 * (size_t)K); down(imp_t, C.imp_t, 2 * (size_t)K);
		down(pen, C.pen, n); down(bounce, C.bounce, n); down(mu, C.mu, n); down(imp_n, C.imp_n, n);
		down(life, C.life, n); down(live, C.live, n); down(sr, C.sready, n);
		std::vector<char> gone = w->removed_contacts(life);
		int o = 0, pos = 0;
		uint64_t prev = ~0ull;
		for (int k = 0; k < n; k++) {
			if (gone[k]) continue;
			if (key[k] != prev) pos = 0;
			prev = key[k];
			if (o < cap && ii && mg && ff) {
				int64_t* r = ii + 6 * (size_t)o;
				r[0] = (int64_t)(key[k] >> 32) + 1;
				r[1] = (int64_t)(key[k] & 0xFFFFFFFFu) + 1;
				r[2] = life[k]; r[3] = live[k]; r[4] = sr[k]; r[5] = pos;
				mg[2 * o] = cid[k];
				mg[2 * o + 1] = bid[k];
				double* f = ff + 21 * (size_t)o;
				for (int c = 0; c < 3; c++) {
					f[c] = lp1[c * (size_t)K + k];
					f[3 + c] = lp2[c * (size_t)K + k];
					f[6 + c] = ln1[c * (size_t)K + k];
					f[15 + c] = fu[c * (size_t)K + k];
					f[18 + c] = fw[c * (size_t)K + k];
				}
				f[9] = pen[k]; f[10] = bounce[k]; f[11] = mu[k]; f[12] = imp_n[k]; f[13] = imp_t[k]; f[14] = imp_t[K + k];
			}
			pos++;
			o++;
		}
		*nout = o;
		if (o > cap) status = CP_ERR_CAPACITY;
	});
	return g != CP_OK ? g : status;
}

int cp_world_set_node_order(cp_world* w, int n, const int64_t* ids) {
	if (!w || n < 0 || (n > 0 && !ids)) return CP_ERR_INVALID;
	for (int k = 0; k < n; k++)
		if (ids[k] < 1 || ids[k] > w->nb) return CP_ERR_INVALID;
	w->node_order.assign(ids, ids + n);
	return CP_OK;
}

int cp_joint_add(cp_world* w, int type, int b1, int b2, const double* geo14, int* joint_index) {
	static const int degree[14] = {0, 1, 0, 3, 4, 5, 6, 5, 1, 1, 1, 1, 1, 0};  // rows per kind (ConstraintSolver.h:193-398)
	if (!w || !geo14 || !joint_index || type < 1 || type > CP_JOINT_SPRING || (degree[type] == 0 && type != CP_JOINT_SPRING))
		return CP_ERR_INVALID;
	if (b1 < 0 || b2 < 0 || b1 >= w->nb || b2 >= w->nb || b1 == b2 || w->b_scene[b1] != w->b_scene[b2]) return CP_ERR_INVALID;
	for (int c = 0; c < 14; c++)
		if (!std::isfinite(geo14[c])) return CP_ERR_INVALID;
	int k = (int)w->j_type.size();
	w->j_type.push_back(degree[type]);
	w->j_kind.push_back(type);
	w->j_a.push_back(b1);
	w->j_b.push_back(b2);
	w->j_geo.insert(w->j_geo.end(), geo14, geo14 + 14);
	w->j_drive.insert(w->j_drive.end(), {0.0, 0.0, 0.0, 0.0});  // powered mode off (the Motor / Piston constructors)
	w->j_cfm.push_back(std::numeric_limits<double>::quiet_NaN());  // CFM{USE_GLOBAL}
	uint64_t key = (uint64_t)std::min(b1, b2) << 32 | (uint64_t)std::max(b1, b2);
	// limits, motors and pistons are not holonomic: they join the pair's edge
	// and leave the holonomic systems as they are (PhysicsEngine.cpp:725-760,
	// :808-846, :926-970)
	const bool holonomic = !joint_is_inequality(type);
	auto it = w->jedges.find(key);
	if (it == w->jedges.end()) {
		cp_world::JEdge e;
		e.n1 = b1;
		e.n2 = b2;
		e.joints.push_back(k);
		e.nhol = holonomic;
		e.scan = holonomic;
		w->jedges.emplace(key, e);
	} else {
		it->second.joints.push_back(k);
		it->second.nhol += holonomic;
		if (holonomic) {
			if (it->second.sys >= 0) w->hs_changed = true;  // constraints_changed_flag (PhysicsEngine.cpp:679-681)
			else it->second.scan = true;
		}
	}
	w->joints_dirty = true;
	if (holonomic) w->hs_scan = true;
	*joint_index = k;
	return CP_OK;
}

int cp_joint_set_drive(cp_world* w, int k, int mode, double max, double target) {
	if (!w || k < 0 || k >= (int)w->j_type.size() || mode < 0 || mode > 3) return CP_ERR_INVALID;
	if (w->j_kind[k] != CP_JOINT_MOTOR && w->j_kind[k] != CP_JOINT_PISTON_DRIVE) return CP_ERR_INVALID;
	if (!std::isfinite(max) || !std::isfinite(target)) return CP_ERR_INVALID;
	int st = guard(w, [&] {
		// setMotor* / setPiston* (PhysicsEngine.cpp:1683-1912): off keeps the
		// parameters, the other modes set the maximum and their target; both
		// bodies receive a waking action
		double* d = &w->j_drive[4 * (size_t)k];
		d[0] = mode;
		if (mode >= 1) d[1] = max;
		if (mode == 2) d[2] = target;
		if (mode == 3) d[3] = target;
		w->drive_dirty = true;
	});
	if (st != CP_OK || !w->cfg.sleeping) return st;
	const int ids[2] = {w->j_a[k], w->j_b[k]};
	return cp_world_wake_bodies(w, 2, ids);
}

int cp_joint_set_cfm(cp_world* w, int k, double cfm) {
	if (!w || k < 0 || k >= (int)w->j_type.size() || (!std::isnan(cfm) && !(cfm >= 0 && std::isfinite(cfm)))) return CP_ERR_INVALID;
	return guard(w, [&] {
		w->j_cfm[k] = cfm;
		w->joints_dirty = true;
	});
}

int cp_world_joint_count(cp_world* w, int* n) {
	if (!w || !n) return CP_ERR_INVALID;
	*n = (int)w->j_type.size();
	return CP_OK;
}

int cp_world_set_joint_state(cp_world* w, const double* s12) {
	if (!w || (!s12 && !w->j_type.empty())) return CP_ERR_INVALID;
	return guard(w, [&] {
		if (w->joints_dirty || w->j_dev_n != (int)w->j_type.size()) w->upload_joints();
		if (w->j_dev_n) w->write_joint_state(s12);
		w->j_init_from = w->j_dev_n;  // the caller's state replaces the constructors'
	});
}

int cp_world_get_joint_state(cp_world* w, double* s12) {
	if (!w || (!s12 && !w->j_type.empty())) return CP_ERR_INVALID;
	return guard(w, [&] {
		if (w->joints_dirty || w->j_dev_n != (int)w->j_type.size()) w->upload_joints();
		w->init_new_joints(w->bodies());
		HIPCHK(hipStreamSynchronize(w->stream));
		if (w->j_dev_n) w->read_joint_state(s12, w->j_dev_n);
	});
}

int cp_world_get_joint_rows(cp_world* w, int k, double* ja, double* jb, double* ita, double* itb, int* n) {
	if (!w || !ja || !jb || !ita || !itb || !n || k < 0 || k >= w->j_dev_n) return CP_ERR_INVALID;
	return guard(w, [&] {
		double R[CP_JRS];
		HIPCHK(hipStreamSynchronize(w->stream));
		HIPCHK(hipMemcpy(R, w->dj_row.p + (size_t)k * CP_JRS, sizeof(R), hipMemcpyDeviceToHost));
		int t = w->j_type[k];
		for (int r = 0; r < t; r++)
			for (int c = 0; c < 6; c++) {
				ja[r * 6 + c] = R[CP_J_JA + r * 6 + c];
				jb[r * 6 + c] = R[CP_J_JB + r * 6 + c];
				ita[c * t + r] = R[CP_J_ITA + c * CP_JMAX + r];
				itb[c * t + r] = R[CP_J_ITB + c * CP_JMAX + r];
			}
		*n = t;
	});
}

int cp_world_set_holonomic_systems(cp_world* w, int n_sys, const int64_t* off, const int64_t* joints) {
	if (!w || n_sys < 0 || (n_sys > 0 && (!off || !joints))) return CP_ERR_INVALID;
	const int K = (int)w->j_type.size();
	std::vector<char> used(K, 0);
	std::vector<std::vector<int>> og(n_sys);
	for (int q = 0; q < n_sys; q++) {
		if (off[q] > off[q + 1] || (q == 0 && off[0] != 0)) return CP_ERR_INVALID;
		for (int64_t i = off[q]; i < off[q + 1]; i++) {
			if (joints[i] < 0 || joints[i] >= K || used[joints[i]]) return CP_ERR_INVALID;
			if (joint_is_inequality(w->j_kind[joints[i]])) return CP_ERR_INVALID;  // inequality kinds are not holonomic
			used[joints[i]] = 1;
			og[q].push_back((int)joints[i]);
		}
	}
	return guard(w, [&] {
		w->hs_og.swap(og);
		w->hs_explicit = true;
		w->layout_systems();
	});
}

int cp_world_holonomic_system_count(cp_world* w, int* n) {
	if (!w || !n) return CP_ERR_INVALID;
	*n = (int)w->hs_og.size();
	return CP_OK;
}

int cp_world_get_holonomic_system(cp_world* w, int k, int64_t* og, int64_t* order, int64_t* table, int64_t* vloc, double* factor,
                                  int64_t* sizes4) {
	if (!w || !sizes4 || k < 0 || k >= (int)w->hs_lay.size()) return CP_ERR_INVALID;
	const cp_world::HLayout& L = w->hs_lay[k];
	sizes4[0] = (int64_t)L.order.size();
	sizes4[1] = (int64_t)L.table.size();
	sizes4[2] = L.degree;
	sizes4[3] = L.cap;
	for (size_t i = 0; og && i < w->hs_og[k].size(); i++) og[i] = w->hs_og[k][i];
	for (size_t i = 0; order && i < L.order.size(); i++) order[i] = L.order[i];
	for (size_t i = 0; table && i < L.table.size(); i++) table[i] = L.table[i];
	for (size_t i = 0; vloc && i < L.vloc.size(); i++) vloc[i] = L.vloc[i];
	if (!factor || L.cap == 0) return CP_OK;
	return guard(w, [&] {
		std::vector<int> bo;
		down(bo, w->dh_buf_off.p, (size_t)k + 1);
		HIPCHK(hipStreamSynchronize(w->stream));
		HIPCHK(hipMemcpy(factor, w->dh_buf.p + bo[k], (size_t)L.cap * sizeof(double), hipMemcpyDeviceToHost));
	});
}

// Iteration order of std::unordered_map<unsigned, T> after inserting 1..n in
// sequence — the container type the reference keeps its constraint-graph nodes
// in (PhysicsEngine.h:230); produced by the same standard library.
int cp_reference_node_order(int n, int64_t* ids) {
	if (n < 0 || (n > 0 && !ids)) return CP_ERR_INVALID;
	std::unordered_map<unsigned int, int> m;
	for (int i = 1; i <= n; i++) m[(unsigned)i] = i;
	int k = 0;
	for (auto& kv : m) ids[k++] = kv.first;
	return CP_OK;
}

int cp_world_save_snapshot(cp_world* w, const char* path) {
	if (!w || !path || !w->uploaded) return CP_ERR_INVALID;
	SnapFile f;
	int nb = w->nb, nj = (int)w->j_type.size(), nc = 0;
	f.state.resize(13 * (size_t)nb);
	f.inertia.resize(18 * (size_t)nb);
	int st = cp_world_get_state(w, f.state.data());
	if (st == CP_OK) st = cp_world_get_inertia(w, f.inertia.data());
	if (st != CP_OK) return st;
	st = cp_world_get_contacts(w, 0, nullptr, nullptr, nullptr, &nc);
	if (st != CP_OK && st != CP_ERR_CAPACITY) return st;
	f.ci.resize(6 * (size_t)nc);
	f.cm.resize(2 * (size_t)nc);
	f.cf.resize(21 * (size_t)nc);
	if (nc && (st = cp_world_get_contacts(w, nc, f.ci.data(), f.cm.data(), f.cf.data(), &nc)) != CP_OK) return st;
	f.joint.resize(CP_JSTATE * (size_t)nj);
	if (nj && (st = cp_world_get_joint_state(w, f.joint.data())) != CP_OK) return st;
	if (nj) {
		// powered modes, and the rows, projections and pseudo impulses an
		// inequality joint solves with when its update leaves them standing
		f.jdrive = w->j_drive;
		st = guard(w, [&] {
			HIPCHK(hipStreamSynchronize(w->stream));
			down(f.jrow, w->dj_row.p, (size_t)CP_JRS * nj);
			down(f.jpimp, w->dj_pimp.p, (size_t)CP_JMAX * nj);
			std::vector<int> pj;
			down(pj, w->dj_proj.p, nj);
			f.jproj.assign(pj.begin(), pj.end());
		});
		if (st != CP_OK) return st;
	}
	f.counts[0] = nb;
	f.counts[1] = nc;
	f.counts[2] = nj;
	if (w->sl_active) {
		// sleep state: per body asleep, asleep at the last removal, ticks,
		// history length and head, sleep disabled; counters; the history ring
		st = guard(w, [&] {
			HIPCHK(hipStreamSynchronize(w->stream));
			const int hc = std::max(w->sl_hcap, 0);
			std::vector<int> a, rm, t, hl, hh;
			down(a, w->sl_asleep.p, nb); down(rm, w->sl_asleep_rm.p, nb); down(t, w->sl_ticks.p, nb);
			down(hl, w->sl_hlen.p, nb); down(hh, w->sl_hhead.p, nb);
			down(f.sl_counter, w->sl_counter.p, nb);
			if (hc) down(f.sl_hist, w->sl_hist.p, 6 * (size_t)hc * nb);
			w->sl_dis_host.resize(nb, 0);
			f.sl_ints.resize(6 * (size_t)nb);
			for (int i = 0; i < nb; i++) {
				int64_t* r = &f.sl_ints[6 * (size_t)i];
				r[0] = a[i]; r[1] = rm[i]; r[2] = t[i]; r[3] = hl[i]; r[4] = hh[i]; r[5] = w->sl_dis_host[i];
			}
			f.counts[3] = hc + 1;
		});
		if (st != CP_OK) return st;
	}
	FILE* fp = fopen(path, "wb");
	if (!fp) return CP_ERR_IO;
	uint64_t h = 14695981039346656037ull;
	bool ok = true;
	auto put = [&](const void* p, size_t bytes) {
		if (!ok || !bytes) return;
		ok = fwrite(p, 1, bytes, fp) == bytes;
		h = fnv1a(h, p, bytes);
	};
	put(SNAP_MAGIC, 8);
	put(f.counts, sizeof f.counts);
	put(f.state.data(), f.state.size() * 8);
	put(f.inertia.data(), f.inertia.size() * 8);
	put(f.ci.data(), f.ci.size() * 8);
	put(f.cm.data(), f.cm.size() * 8);
	put(f.cf.data(), f.cf.size() * 8);
	put(f.joint.data(), f.joint.size() * 8);
	put(f.jdrive.data(), f.jdrive.size() * 8);
	put(f.jproj.data(), f.jproj.size() * 8);
	put(f.jrow.data(), f.jrow.size() * 8);
	put(f.jpimp.data(), f.jpimp.size() * 8);
	put(f.sl_ints.data(), f.sl_ints.size() * 8);
	put(f.sl_counter.data(), f.sl_counter.size() * 8);
	put(f.sl_hist.data(), f.sl_hist.size() * 8);
	ok = ok && fwrite(&h, 8, 1, fp) == 1;
	ok = (fclose(fp) == 0) && ok;
	return ok ? CP_OK : CP_ERR_IO;
}

int cp_world_load_snapshot(cp_world* w, const char* path) {
	if (!w || !path || w->nb == 0) return CP_ERR_INVALID;
	FILE* fp = fopen(path, "rb");
	if (!fp) return CP_ERR_IO;
	SnapFile f;
	uint64_t h = 14695981039346656037ull;
	bool ok = true;
	auto get = [&](void* p, size_t bytes) {
		if (!ok || !bytes) return;
		ok = fread(p, 1, bytes, fp) == bytes;
		if (ok) h = fnv1a(h, p, bytes);
	};
	char magic[8] = {};
	get(magic, 8);
	ok = ok && memcmp(magic, SNAP_MAGIC, 8) == 0;
	get(f.counts, sizeof f.counts);
	ok = ok && f.counts[0] == w->nb && f.counts[2] == (int64_t)w->j_type.size() && f.counts[1] >= 0 && f.counts[1] < (int64_t)1 << 31 &&
	     f.counts[3] >= 0 && f.counts[3] <= 100001;
	if (ok) {
		size_t nb = (size_t)f.counts[0], nc = (size_t)f.counts[1], nj = (size_t)f.counts[2];
		f.state.resize(13 * nb); f.inertia.resize(18 * nb);
		f.ci.resize(6 * nc); f.cm.resize(2 * nc); f.cf.resize(21 * nc);
		f.joint.resize(CP_JSTATE * nj);
		f.jdrive.resize(4 * nj); f.jproj.resize(nj); f.jrow.resize((size_t)CP_JRS * nj); f.jpimp.resize((size_t)CP_JMAX * nj);
		get(f.state.data(), f.state.size() * 8);
		get(f.inertia.data(), f.inertia.size() * 8);
		get(f.ci.data(), f.ci.size() * 8);
		get(f.cm.data(), f.cm.size() * 8);
		get(f.cf.data(), f.cf.size() * 8);
		get(f.joint.data(), f.joint.size() * 8);
		get(f.jdrive.data(), f.jdrive.size() * 8);
		get(f.jproj.data(), f.jproj.size() * 8);
		get(f.jrow.data(), f.jrow.size() * 8);
		get(f.jpimp.data(), f.jpimp.size() * 8);
		if (f.counts[3] > 0) {
			f.sl_ints.resize(6 * nb);
			f.sl_counter.resize(nb);
			f.sl_hist.resize(6 * (size_t)(f.counts[3] - 1) * nb);
			get(f.sl_ints.data(), f.sl_ints.size() * 8);
			get(f.sl_counter.data(), f.sl_counter.size() * 8);
			get(f.sl_hist.data(), f.sl_hist.size() * 8);
		}
		uint64_t want = 0;
		ok = ok && fread(&want, 8, 1, fp) == 1 && want == h;
	}
	fclose(fp);
	if (!ok) return CP_ERR_IO;
	int st = cp_world_set_state(w, f.state.data(), f.inertia.data());
	if (st == CP_OK) st = cp_world_set_contacts(w, (int)f.counts[1], f.ci.data(), f.cm.data(), f.cf.data());
	if (st == CP_OK && f.counts[2] > 0) st = cp_world_set_joint_state(w, f.joint.data());
	if (st == CP_OK && f.counts[2] > 0)
		st = guard(w, [&] {
			w->j_drive = f.jdrive;
			w->upload_drives();
			std::vector<int> pj(f.jproj.begin(), f.jproj.end());
			up(w->dj_proj, pj);
			up(w->dj_pimp, f.jpimp);
			HIPCHK(hipMemcpy(w->dj_row.p, f.jrow.data(), f.jrow.size() * 8, hipMemcpyHostToDevice));
		});
	if (st == CP_OK && f.counts[3] > 0)
		st = guard(w, [&] {
			const int nb = w->nb, hc = (int)f.counts[3] - 1;
			std::vector<int> a(nb), rm(nb), t(nb), hl(nb), hh(nb);
			w->sl_dis_host.assign(nb, 0);
			for (int i = 0; i < nb; i++) {
				const int64_t* r = &f.sl_ints[6 * (size_t)i];
				a[i] = (int)r[0]; rm[i] = (int)r[1]; t[i] = (int)r[2]; hl[i] = (int)r[3]; hh[i] = (int)r[4];
				w->sl_dis_host[i] = (int)r[5];
			}
			up(w->sl_asleep, a); up(w->sl_asleep_rm, rm); up(w->sl_ticks, t); up(w->sl_hlen, hl); up(w->sl_hhead, hh);
			up(w->sl_disabled, w->sl_dis_host); up(w->sl_counter, f.sl_counter);
			w->sl_hist.reserve(6 * (size_t)std::max(hc, 1) * nb);
			if (hc) HIPCHK(hipMemcpy(w->sl_hist.p, f.sl_hist.data(), f.sl_hist.size() * 8, hipMemcpyHostToDevice));
			w->sl_hcap = hc;
			w->sl_active = true;
			w->sl_wake.clear();
		});
	return st;
}

int cp_world_trajectory_open(cp_world* w, const char* path) {
	if (!w || !path) return CP_ERR_INVALID;
	if (w->traj) fclose(w->traj);
	w->traj = fopen(path, "wb");
	return w->traj ? CP_OK : CP_ERR_IO;
}

int cp_world_trajectory_close(cp_world* w) {
	if (!w) return CP_ERR_INVALID;
	int st = CP_OK;
	if (w->traj && fclose(w->traj) != 0) st = CP_ERR_IO;
	w->traj = nullptr;
	return st;
}

int cp_trajectory_frames(const char* path, int n_bodies, int* n_frames) {
	if (!path || n_bodies < 1 || !n_frames) return CP_ERR_INVALID;
	FILE* fp = fopen(path, "rb");
	if (!fp) return CP_ERR_IO;
	bool ok = fseek(fp, 0, SEEK_END) == 0;
	long bytes = ok ? ftell(fp) : -1;
	fclose(fp);
	const long frame = 56L * n_bodies;
	if (bytes < 0 || bytes % frame != 0) return CP_ERR_IO;
	*n_frames = (int)(bytes / frame);
	return CP_OK;
}

int cp_trajectory_read(const char* path, int n_bodies, int frame, double* poses7) {
	if (!path || n_bodies < 1 || frame < 0 || !poses7) return CP_ERR_INVALID;
	int nf = 0;
	int st = cp_trajectory_frames(path, n_bodies, &nf);
	if (st != CP_OK) return st;
	if (frame >= nf) return CP_ERR_INVALID;
	FILE* fp = fopen(path, "rb");
	if (!fp) return CP_ERR_IO;
	bool ok = fseek(fp, (long)frame * 56L * n_bodies, SEEK_SET) == 0 && fread(poses7, 8, 7 * (size_t)n_bodies, fp) == 7 * (size_t)n_bodies;
	fclose(fp);
	return ok ? CP_OK : CP_ERR_IO;
}

int cp_world_serving_init(cp_world* w, int capacity) {
	if (!w || capacity < 1) return CP_ERR_INVALID;
	if (w->nb != 0 || !w->j_type.empty() || w->serving) return CP_ERR_STATE;
	return guard(w, [&] {
		double c[3] = {0, 0, 0};
		w->park_shape = w->hs.add_sphere(c, 0.5);
		w->nb = capacity;
		w->b_shape.assign(capacity, w->park_shape);
		w->b_move.assign(capacity, 1);
		w->b_scene.assign(capacity, 0);
		w->b_massprops.assign(19 * (size_t)capacity, 0.0);
		w->b_mat.assign(4 * (size_t)capacity, 0.0);
		w->serving = true;
		w->upload();
		for (size_t& d : w->sh_done) d = 0;
		w->ranges = {{0, capacity, -1}};
		w->park_range(0, capacity);
		w->serving_dirty = true;
		w->serving_commit();
	});
}

int cp_world_admit_scene(cp_world* w, int n, const int* shape_ids, const int* movement, const double* mp, const double* material,
                         const double* state13, int* handle) {
	if (!w || !w->serving || n < 1 || !shape_ids || !movement || !mp || !material || !state13 || !handle) return CP_ERR_INVALID;
	for (int i = 0; i < n; i++) {
		int sid = shape_ids[i];
		if (sid < 0 || sid >= (int)w->hs.type.size() || (movement[i] != 0 && movement[i] != 1)) return CP_ERR_INVALID;
		if (w->hs.type[sid] == SHAPE_MESH) return CP_ERR_INVALID;
		if (w->hs.pt_off[sid + 1] - w->hs.pt_off[sid] > SLOT_PTS || w->hs.face_off[sid + 1] - w->hs.face_off[sid] > SLOT_NRM)
			return CP_ERR_INVALID;  // larger than a serving slot
	}
	int r = -1;
	for (int k = 0; k < (int)w->ranges.size(); k++)
		if (w->ranges[k].handle < 0 && w->ranges[k].len >= n) { r = k; break; }
	if (r < 0) return CP_ERR_CAPACITY;  // no free range this large: the caller queues the request
	cp_world::Range fr = w->ranges[r];
	int h = w->next_handle++;
	w->ranges[r] = {fr.start, n, h};
	if (fr.len > n) w->ranges.insert(w->ranges.begin() + r + 1, cp_world::Range{fr.start + n, fr.len - n, -1});
	for (int i = 0; i < n; i++)
		w->stage_body(fr.start + i, shape_ids[i], movement[i], mp + 19 * (size_t)i, material + 4 * (size_t)i, state13 + 13 * (size_t)i);
	w->serving_dirty = true;
	*handle = h;
	return CP_OK;
}

static int find_range(cp_world* w, int handle) {
	for (int k = 0; k < (int)w->ranges.size(); k++)
		if (w->ranges[k].handle == handle) return k;
	return -1;
}

int cp_world_retire_scene(cp_world* w, int handle) {
	if (!w || !w->serving) return CP_ERR_INVALID;
	int r = find_range(w, handle);
	if (r < 0) return CP_ERR_INVALID;
	cp_world::Range& R = w->ranges[r];
	w->park_range(R.start, R.len);
	w->kill_ranges.insert(w->kill_ranges.end(), {R.start, R.len});
	R.handle = -1;
	// merge free neighbours
	if (r + 1 < (int)w->ranges.size() && w->ranges[r + 1].handle < 0) {
		w->ranges[r].len += w->ranges[r + 1].len;
		w->ranges.erase(w->ranges.begin() + r + 1);
	}
	if (r > 0 && w->ranges[r - 1].handle < 0) {
		w->ranges[r - 1].len += w->ranges[r].len;
		w->ranges.erase(w->ranges.begin() + r);
	}
	w->serving_dirty = true;
	return CP_OK;
}

int cp_world_serving_commit(cp_world* w) {
	if (!w || !w->serving) return CP_ERR_INVALID;
	return guard(w, [&] { w->serving_commit(); });
}

int cp_world_get_scene_state(cp_world* w, int handle, double* state13, int* n) {
	if (!w || !w->serving || !n) return CP_ERR_INVALID;
	int r = find_range(w, handle);
	if (r < 0) return CP_ERR_INVALID;
	return guard(w, [&] {
		w->serving_commit();
		HIPCHK(hipStreamSynchronize(w->stream));
		const cp_world::Range& R = w->ranges[r];
		*n = R.len;
		if (!state13) return;
		int N = w->nb;
		std::vector<double> buf((size_t)R.len);
		auto get = [&](DArr<double>& a, int comps, int col) {
			for (int c = 0; c < comps; c++) {
				HIPCHK(hipMemcpy(buf.data(), a.p + (size_t)c * N + R.start, R.len * sizeof(double), hipMemcpyDeviceToHost));
				for (int i = 0; i < R.len; i++) state13[13 * (size_t)i + col + c] = buf[i];
			}
		};
		get(w->com, 3, 0);
		get(w->q, 4, 3);
		get(w->vel, 3, 7);
		get(w->angv, 3, 10);
		std::vector<int> f;
		f.resize((size_t)R.len);
		HIPCHK(hipMemcpy(f.data(), w->nf_body.p + R.start, R.len * sizeof(int), hipMemcpyDeviceToHost));
		for (int v : f)
			if (v) throw ApiErr{CP_ERR_NONFINITE};  // the state is returned; the status names the divergence
	});
}

int cp_world_get_scene_states(cp_world* w, int n, const int* handles, int cap_rows, double* state13, int* counts, int* status) {
	if (!w || !w->serving || n < 0 || (n > 0 && (!handles || !counts || !status)) || cap_rows < 0) return CP_ERR_INVALID;
	std::vector<int> start(n), off(n + 1, 0);
	for (int k = 0; k < n; k++) {
		const int r = find_range(w, handles[k]);
		if (r < 0) return CP_ERR_INVALID;
		start[k] = w->ranges[r].start;
		counts[k] = w->ranges[r].len;
		off[k + 1] = off[k] + counts[k];
	}
	if (off[n] > cap_rows || (off[n] > 0 && !state13)) return CP_ERR_CAPACITY;
	return guard(w, [&] {
		w->serving_commit();
		if (!n) return;
		// one gather on the device, one copy back
		w->gs_idx.reserve(2 * (size_t)n + 1);
		std::vector<int> idx(start);
		idx.insert(idx.end(), off.begin(), off.end());
		HIPCHK(hipMemcpyAsync(w->gs_idx.p, idx.data(), idx.size() * sizeof(int), hipMemcpyHostToDevice, w->stream));
		w->gs_rows.reserve(13 * (size_t)std::max(off[n], 1));
		w->gs_flag.reserve((size_t)n);
		HIPCHK(hipMemsetAsync(w->gs_flag.p, 0, (size_t)n * sizeof(int), w->stream));
		launch_gather_states(w->bodies(), w->gs_idx.p, w->gs_idx.p + n, n, off[n], w->nf_body.p, w->gs_rows.p, w->gs_flag.p, w->stream);
		HIPCHK(hipMemcpyAsync(state13, w->gs_rows.p, 13 * (size_t)off[n] * sizeof(double), hipMemcpyDeviceToHost, w->stream));
		HIPCHK(hipMemcpyAsync(status, w->gs_flag.p, (size_t)n * sizeof(int), hipMemcpyDeviceToHost, w->stream));
		HIPCHK(hipStreamSynchronize(w->stream));
		for (int k = 0; k < n; k++) status[k] = status[k] ? CP_ERR_NONFINITE : CP_OK;
	});
}

// info4: capacity, free bodies, largest free range, admitted scenes.
int cp_world_scene_snapshot(cp_world* w, int handle, int cap_contacts, double* state13, double* inertia18, int64_t* ci6, uint64_t* cm2,
                            double* cf21, int* n_bodies, int* n_contacts) {
	if (!w || !w->serving || !n_bodies || !n_contacts) return CP_ERR_INVALID;
	int r = find_range(w, handle);
	if (r < 0) return CP_ERR_INVALID;
	int st = guard(w, [&] { w->serving_commit(); });
	if (st != CP_OK) return st;
	const cp_world::Range R = w->ranges[r];
	*n_bodies = R.len;
	int nc = 0;
	st = cp_world_get_contacts(w, 0, nullptr, nullptr, nullptr, &nc);
	if (st != CP_OK && st != CP_ERR_CAPACITY) return st;
	std::vector<int64_t> ci(6 * (size_t)nc);
	std::vector<uint64_t> cm(2 * (size_t)nc);
	std::vector<double> cf(21 * (size_t)nc);
	if (nc && (st = cp_world_get_contacts(w, nc, ci.data(), cm.data(), cf.data(), &nc)) != CP_OK) return st;
	const int64_t lo = R.start, hi = (int64_t)R.start + R.len;  // ids (lo, hi]
	const uint64_t shift = 16 * (uint64_t)R.start;
	int m = 0;
	for (int k = 0; k < nc; k++) {
		if (ci[6 * (size_t)k] <= lo || ci[6 * (size_t)k] > hi) continue;
		if (state13 && m < cap_contacts && ci6 && cm2 && cf21) {
			std::copy_n(&ci[6 * (size_t)k], 6, ci6 + 6 * (size_t)m);
			ci6[6 * (size_t)m] -= lo;
			ci6[6 * (size_t)m + 1] -= lo;
			uint64_t c = cm[2 * (size_t)k];
			uint64_t clo = (c & 0xFFFFFFFFull) - shift, chi = (c >> 32) - shift;
			cm2[2 * (size_t)m] = (clo & 0xFFFFFFFFull) | (chi << 32);
			cm2[2 * (size_t)m + 1] = cm[2 * (size_t)k + 1];
			std::copy_n(&cf[21 * (size_t)k], 21, cf21 + 21 * (size_t)m);
		}
		m++;
	}
	*n_contacts = m;
	if (!state13) return CP_OK;
	if (m > cap_contacts) return CP_ERR_CAPACITY;
	int n = 0;
	st = cp_world_get_scene_state(w, handle, state13, &n);
	if (st != CP_OK && st != CP_ERR_NONFINITE) return st;
	int st2 = CP_OK;
	if (inertia18)
		st2 = guard(w, [&] {
			const size_t N = (size_t)w->nb;
			std::vector<double> comp((size_t)R.len);
			for (int c = 0; c < 18; c++) {
				const double* src = (c < 9 ? w->I.p + c * N : w->Iinv.p + (c - 9) * N) + R.start;
				HIPCHK(hipMemcpy(comp.data(), src, (size_t)R.len * sizeof(double), hipMemcpyDeviceToHost));
				for (int i = 0; i < R.len; i++) inertia18[18 * (size_t)i + c] = comp[i];
			}
		});
	return st2 != CP_OK ? st2 : st;
}

int cp_world_admit_scene_resume(cp_world* w, int n, const int* shape_ids, const int* movement, const double* mp, const double* material,
                                const double* state13, const double* inertia18, int n_contacts, const int64_t* ci6, const uint64_t* cm2,
                                const double* cf21, int* handle) {
	if (!w || !w->serving || !inertia18 || n_contacts < 0 || (n_contacts > 0 && (!ci6 || !cm2 || !cf21)) || !handle) return CP_ERR_INVALID;
	for (int k = 0; k < n_contacts; k++) {
		const int64_t* r = ci6 + 6 * (size_t)k;
		if (r[0] < 1 || r[1] <= r[0] || r[1] > n) return CP_ERR_INVALID;
		if (k > 0) {
			const int64_t* p = ci6 + 6 * (size_t)(k - 1);
			if (p[0] > r[0] || (p[0] == r[0] && p[1] > r[1])) return CP_ERR_INVALID;  // grouped by sorted pair
		}
	}
	int st = cp_world_admit_scene(w, n, shape_ids, movement, mp, material, state13, handle);
	if (st != CP_OK) return st;
	int r = find_range(w, *handle);
	cp_world::Resume res;
	res.start = w->ranges[r].start;
	res.n = n;
	res.inertia.assign(inertia18, inertia18 + 18 * (size_t)n);
	res.ci.assign(ci6, ci6 + 6 * (size_t)n_contacts);
	res.cm.assign(cm2, cm2 + 2 * (size_t)n_contacts);
	res.cf.assign(cf21, cf21 + 21 * (size_t)n_contacts);
	const uint64_t shift = 16 * (uint64_t)res.start;
	for (int k = 0; k < n_contacts; k++) {
		res.ci[6 * (size_t)k] += res.start;
		res.ci[6 * (size_t)k + 1] += res.start;
		uint64_t c = res.cm[2 * (size_t)k];
		uint64_t clo = (c & 0xFFFFFFFFull) + shift, chi = (c >> 32) + shift;
		res.cm[2 * (size_t)k] = (clo & 0xFFFFFFFFull) | (chi << 32);
	}
	w->resumes.push_back(std::move(res));
	return CP_OK;
}

int cp_world_serving_info(cp_world* w, int* info4) {
	if (!w || !w->serving || !info4) return CP_ERR_INVALID;
	int fr = 0, big = 0, act = 0;
	for (const cp_world::Range& r : w->ranges) {
		if (r.handle < 0) {
			fr += r.len;
			big = std::max(big, r.len);
		} else act++;
	}
	info4[0] = w->nb;
	info4[1] = fr;
	info4[2] = big;
	info4[3] = act;
	return CP_OK;
}

int cp_world_step(cp_world* w, int n_steps) {
	if (!w || n_steps < 0 || w->nb == 0) return CP_ERR_INVALID;
	if (!w->uploaded) return CP_ERR_STATE;
	int st = guard(w, [&] {
		if (!w->nf_host) HIPCHK(hipHostMalloc((void**)&w->nf_host, sizeof(int), hipHostMallocDefault));
		launch_fill_int(w->nf_any.p, 1, 0, w->stream);
		for (int i = 0; i < n_steps; i++) {
			w->step();
			if (w->traj) w->record_frame();
		}
		HIPCHK(hipMemcpyAsync(w->nf_host, w->nf_any.p, sizeof(int), hipMemcpyDeviceToHost, w->stream));
		HIPCHK(hipStreamSynchronize(w->stream));
		w->nf_step = *w->nf_host != 0;
	});
	if (st == CP_OK && w->nf_step) return CP_ERR_NONFINITE;
	return st;
}

int cp_world_scene_flags(cp_world* w, int cap, int* flags, int* n) {
	if (!w || !n || !w->uploaded) return CP_ERR_INVALID;
	return guard(w, [&] {
		*n = w->nscenes;
		if (w->nscenes > cap || !flags) throw ApiErr{CP_ERR_CAPACITY};
		HIPCHK(hipStreamSynchronize(w->stream));
		std::vector<int> f, sbo;
		down(f, w->nf_body.p, w->nb);
		down(sbo, w->scene_body_off.p, (size_t)w->nscenes + 1);
		for (int s2 = 0; s2 < w->nscenes; s2++) {
			flags[s2] = 0;
			for (int i = sbo[s2]; i < sbo[s2 + 1]; i++) flags[s2] |= f[i];
		}
	});
}

int cp_world_get_sleep_state(cp_world* w, double* out4) {
	if (!w || !out4) return CP_ERR_INVALID;
	if (!w->uploaded) return CP_ERR_STATE;
	return guard(w, [&] {
		HIPCHK(hipStreamSynchronize(w->stream));
		std::vector<int> a, t, h;
		std::vector<double> c;
		down(a, w->sl_asleep.p, w->nb); down(t, w->sl_ticks.p, w->nb); down(h, w->sl_hlen.p, w->nb); down(c, w->sl_counter.p, w->nb);
		for (int i = 0; i < w->nb; i++) {
			out4[4 * (size_t)i] = a[i] && w->b_move[i] == 0 ? 1.0 : 0.0;
			out4[4 * (size_t)i + 1] = c[i];
			out4[4 * (size_t)i + 2] = (double)t[i];
			out4[4 * (size_t)i + 3] = (double)h[i];
		}
	});
}

int cp_world_wake_bodies(cp_world* w, int n, const int* idx) {
	if (!w || n < 0 || (n > 0 && !idx)) return CP_ERR_INVALID;
	for (int k = 0; k < n; k++)
		if (idx[k] < 0 || idx[k] >= w->nb) return CP_ERR_INVALID;
	return guard(w, [&] {
		w->sl_wake.insert(w->sl_wake.end(), idx, idx + n);
		if (!w->uploaded || !n) return;
		// RigidBody::alertWakingAction (RigidBody.cpp:340-345): counters and
		// history reset at once
		HIPCHK(hipStreamSynchronize(w->stream));
		std::vector<int> t, h;
		std::vector<double> c;
		down(t, w->sl_ticks.p, w->nb); down(h, w->sl_hlen.p, w->nb); down(c, w->sl_counter.p, w->nb);
		for (int k = 0; k < n; k++) {
			t[idx[k]] = 0;
			h[idx[k]] = 0;
			c[idx[k]] = 0;
		}
		up(w->sl_ticks, t); up(w->sl_hlen, h); up(w->sl_counter, c);
	});
}

int cp_world_set_sleep_disabled(cp_world* w, int n, const int* idx, int disabled) {
	if (!w || n < 0 || (n > 0 && !idx)) return CP_ERR_INVALID;
	for (int k = 0; k < n; k++)
		if (idx[k] < 0 || idx[k] >= w->nb) return CP_ERR_INVALID;
	return guard(w, [&] {
		w->sl_dis_host.resize(w->nb, 0);
		for (int k = 0; k < n; k++) w->sl_dis_host[idx[k]] = disabled ? 1 : 0;
		if (w->uploaded) up(w->sl_disabled, w->sl_dis_host);
	});
}

int cp_broadphase(cp_world* w, int cap, int64_t* out, int* n) {
	if (!w || !n || !w->uploaded) return CP_ERR_INVALID;
	int status = CP_OK;
	int g = guard(w, [&] {
		int np_ = w->broadphase();
		std::vector<uint64_t> h;
		down(h, w->pairs.p, np_);
		*n = np_;
		if (np_ > cap || !out) {
			status = CP_ERR_CAPACITY;
			return;
		}
		for (int k = 0; k < np_; k++) {
			out[2 * k] = (int64_t)(h[k] >> 32) + 1;
			out[2 * k + 1] = (int64_t)(h[k] & 0xFFFFFFFFu) + 1;
		}
	});
	return g != CP_OK ? g : status;
}

// Runs the narrowphase on the given pairs; returns per slot the manifold and
// the input pair it belongs to (a pair with a static mesh owns several slots).
static void narrowphase_slots(cp_world* w, int n, const int64_t* pairs2, std::vector<int>& pair_of, std::vector<int>& cnt,
                              std::vector<double>& nrm, std::vector<double>& mp, std::vector<double>& pp, std::vector<uint64_t>& mg) {
	std::vector<uint64_t> keys(n);
	for (int k = 0; k < n; k++) keys[k] = ((uint64_t)(pairs2[2 * k] - 1) << 32) | (uint64_t)(pairs2[2 * k + 1] - 1);
	up(w->pairs, keys);
	int ns = w->narrowphase(n);
	HIPCHK(hipStreamSynchronize(w->stream));
	int K = w->mcap;
	down(cnt, w->man.count, ns);
	down(nrm, w->man.normal, 3 * (size_t)K);
	down(mp, w->man.maxpen, ns);
	down(pp, w->man.pts, 16 * (size_t)ns);
	down(mg, w->man.magic, 8 * (size_t)ns);
	pair_of.resize(ns);
	if (w->last_expanded) {
		std::vector<int> rank;
		down(rank, w->pair_rank.p, ns);
		int k = -1;
		for (int s = 0; s < ns; s++) {
			if (rank[s] == 0) k++;
			pair_of[s] = k;
		}
	} else {
		for (int s = 0; s < ns; s++) pair_of[s] = s;
	}
}

int cp_narrowphase(cp_world* w, int n, const int64_t* pairs2, int* counts, double* heads, double* pts, uint64_t* magic) {
	if (!w || n < 0 || !w->uploaded || (n > 0 && (!pairs2 || !counts || !heads || !pts || !magic))) return CP_ERR_INVALID;
	for (int k = 0; k < n; k++)
		if (pairs2[2 * k] < 1 || pairs2[2 * k + 1] <= pairs2[2 * k] || pairs2[2 * k + 1] > w->nb) return CP_ERR_INVALID;
	return guard(w, [&] {
		std::vector<int> pair_of, cnt;
		std::vector<double> nrm, mp, pp;
		std::vector<uint64_t> mg;
		narrowphase_slots(w, n, pairs2, pair_of, cnt, nrm, mp, pp, mg);
		int K = w->mcap;
		std::vector<char> done(n, 0);
		for (int s = 0; s < (int)pair_of.size(); s++) {
			int k = pair_of[s];
			if (done[k]) continue;  // first manifold of each pair
			done[k] = 1;
			counts[k] = cnt[s];
			for (int c = 0; c < 3; c++) heads[4 * k + c] = cnt[s] ? nrm[c * (size_t)K + s] : 0;
			heads[4 * k + 3] = mp[s];
			for (int j = 0; j < 16; j++) pts[16 * (size_t)k + j] = pp[16 * (size_t)s + j];
			for (int j = 0; j < 8; j++) magic[8 * (size_t)k + j] = mg[8 * (size_t)s + j];
		}
	});
}

int cp_narrowphase_noalloc(cp_world* w, int n, const int64_t* pairs2, int* counts, double* heads, double* pts, uint64_t* magic) {
	if (!w || n < 0 || !w->uploaded || (n > 0 && (!pairs2 || !counts || !heads || !pts || !magic))) return CP_ERR_INVALID;
	if (n > NA_MAX_PAIRS) return CP_ERR_CAPACITY;
	if (!w->mesh_body.empty()) return CP_ERR_STATE;  // mesh pairs expand into several slots each
	for (int k = 0; k < n; k++)
		if (pairs2[2 * k] < 1 || pairs2[2 * k + 1] <= pairs2[2 * k] || pairs2[2 * k + 1] > w->nb) return CP_ERR_INVALID;
	return guard(w, [&] {
		if (n == 0) return;
		if (w->device < 0 || w->device >= NA_MAX_DEVICES || !na_ptr[w->device]) throw ApiErr{CP_ERR_STATE};
		unsigned char* base = static_cast<unsigned char*>(na_ptr[w->device]);
		hipStream_t st = w->stream;
		uint64_t* keys = reinterpret_cast<uint64_t*>(base + NA_PAIRS);
		double* hd = reinterpret_cast<double*>(base + NA_HEADS);
		HIPCHK(hipMemcpyAsync(hd, pairs2, 16 * (size_t)n, hipMemcpyHostToDevice, st));
		hipLaunchKernelGGL(k_na_keys, dim3((n + 255) / 256), dim3(256), 0, st, reinterpret_cast<const int64_t*>(hd), n, keys);
		Manifolds m{reinterpret_cast<int*>(base + NA_COUNT), reinterpret_cast<double*>(base + NA_NORMAL),
		            reinterpret_cast<double*>(base + NA_MAXPEN), reinterpret_cast<double*>(base + NA_PTS),
		            reinterpret_cast<uint64_t*>(base + NA_MAGIC), reinterpret_cast<double*>(base + NA_MAT)};
		launch_narrowphase(w->bodies(), w->shapes(), w->params(), keys, n, m, NA_MAX_PAIRS, reinterpret_cast<int*>(base + NA_LISTS),
		                   reinterpret_cast<int*>(base + NA_CLASSCNT), w->nocol.p, w->n_nocol, nullptr, st, &w->fk);
		hipLaunchKernelGGL(k_na_heads, dim3((n + 255) / 256), dim3(256), 0, st, m.count, m.normal, m.maxpen, NA_MAX_PAIRS, n, hd);
		HIPCHK(hipMemcpyAsync(counts, m.count, 4 * (size_t)n, hipMemcpyDeviceToHost, st));
		HIPCHK(hipMemcpyAsync(heads, hd, 32 * (size_t)n, hipMemcpyDeviceToHost, st));
		HIPCHK(hipMemcpyAsync(pts, m.pts, 128 * (size_t)n, hipMemcpyDeviceToHost, st));
		HIPCHK(hipMemcpyAsync(magic, m.magic, 64 * (size_t)n, hipMemcpyDeviceToHost, st));
		HIPCHK(hipStreamSynchronize(st));
	});
}

int cp_narrowphase_multi(cp_world* w, int n, const int64_t* pairs2, int cap, int64_t* pair_index, int* counts, double* heads, double* pts,
                         uint64_t* magic, int* n_out) {
	if (!w || n < 0 || cap < 0 || !n_out || !w->uploaded || (n > 0 && !pairs2) || (cap > 0 && (!pair_index || !counts || !heads || !pts || !magic)))
		return CP_ERR_INVALID;
	for (int k = 0; k < n; k++)
		if (pairs2[2 * k] < 1 || pairs2[2 * k + 1] <= pairs2[2 * k] || pairs2[2 * k + 1] > w->nb) return CP_ERR_INVALID;
	int st = guard(w, [&] {
		std::vector<int> pair_of, cnt;
		std::vector<double> nrm, mp, pp;
		std::vector<uint64_t> mg;
		narrowphase_slots(w, n, pairs2, pair_of, cnt, nrm, mp, pp, mg);
		int K = w->mcap, o = 0;
		for (int s = 0; s < (int)pair_of.size(); s++) {
			if (cnt[s] == 0) continue;
			if (o < cap) {
				pair_index[o] = pair_of[s];
				counts[o] = cnt[s];
				for (int c = 0; c < 3; c++) heads[4 * o + c] = nrm[c * (size_t)K + s];
				heads[4 * o + 3] = mp[s];
				for (int j = 0; j < 16; j++) pts[16 * (size_t)o + j] = pp[16 * (size_t)s + j];
				for (int j = 0; j < 8; j++) magic[8 * (size_t)o + j] = mg[8 * (size_t)s + j];
			}
			o++;
		}
		*n_out = o;
	});
	if (st == CP_OK && *n_out > cap) return CP_ERR_CAPACITY;
	return st;
}

int cp_world_last_islands(cp_world* w, int cap_rows, int64_t* rows2, int cap_offs, int64_t* offs, int* n_rows, int* n_islands) {
	if (!w || !n_rows || !n_islands) return CP_ERR_INVALID;
	int st = guard(w, [&] { w->fetch_islands(); });
	if (st != CP_OK) return st;
	int nr = (int)w->last_rows.size() / 2, no = (int)w->last_offs.size();
	*n_rows = nr;
	*n_islands = no - 1;
	if (nr > cap_rows || no > cap_offs || !rows2 || !offs) return CP_ERR_CAPACITY;
	memcpy(rows2, w->last_rows.data(), w->last_rows.size() * sizeof(int64_t));
	memcpy(offs, w->last_offs.data(), w->last_offs.size() * sizeof(int64_t));
	return CP_OK;
}

int cp_world_last_scene_contacts(cp_world* w, int cap, int64_t* first, int* n) {
	if (!w || !n) return CP_ERR_INVALID;
	return guard(w, [&] {
		*n = (int)w->last_scene_coff.size();
		if (*n > cap || !first) throw ApiErr{CP_ERR_CAPACITY};
		std::copy(w->last_scene_coff.begin(), w->last_scene_coff.end(), first);
	});
}

int cp_world_raycast(cp_world* w, int n, const double* origins3, const double* dirs3, const int* scene_of_ray, int* hit_body, double* hit7) {
	if (!w || n < 0 || !w->uploaded || (n > 0 && (!origins3 || !dirs3 || !hit_body || !hit7))) return CP_ERR_INVALID;
	const int ns = w->serving ? w->live_scenes : w->nscenes;
	for (int k = 0; k < n; k++) {
		if (scene_of_ray && (scene_of_ray[k] < -1 || scene_of_ray[k] >= ns)) return CP_ERR_INVALID;
		for (int c = 0; c < 3; c++)
			if (!std::isfinite(origins3[3 * k + c]) || !std::isfinite(dirs3[3 * k + c])) return CP_ERR_INVALID;
	}
	if (n == 0) return CP_OK;
	if (w->mesh_body.size() && !w->meshes_placed) return CP_ERR_STATE;
	return guard(w, [&] {
		w->serving_commit();
		const size_t N = (size_t)n;
		w->ray_io.reserve(13 * N);
		w->ray_int.reserve(2 * N);
		double* org = w->ray_io.p;
		double* dir = org + 3 * N;
		double* out = dir + 3 * N;
		int* sc = w->ray_int.p;
		int* hb = sc + N;
		hipStream_t st = w->stream;
		HIPCHK(hipMemcpyAsync(org, origins3, 24 * N, hipMemcpyHostToDevice, st));
		HIPCHK(hipMemcpyAsync(dir, dirs3, 24 * N, hipMemcpyHostToDevice, st));
		if (scene_of_ray) HIPCHK(hipMemcpyAsync(sc, scene_of_ray, 4 * N, hipMemcpyHostToDevice, st));
		launch_raycast(w->bodies(), w->shapes(), w->meshdev(), w->scene_body_off.p, n, org, dir, scene_of_ray ? sc : nullptr, hb, out, st);
		HIPCHK(hipMemcpyAsync(hit_body, hb, 4 * N, hipMemcpyDeviceToHost, st));
		HIPCHK(hipMemcpyAsync(hit7, out, 56 * N, hipMemcpyDeviceToHost, st));
		HIPCHK(hipStreamSynchronize(st));
		for (int k = 0; k < n; k++)
			if (hit_body[k] < 0)
				for (int c = 0; c < 7; c++) hit7[7 * (size_t)k + c] = 0;
	});
}

int cp_world_last_residual(cp_world* w, double* out) {
	if (!w || !out) return CP_ERR_INVALID;
	if (!(w->cfg.diagnostics & 1)) return CP_ERR_STATE;
	for (int i = 0; i < 4; i++) out[i] = w->last_resid[i];
	return CP_OK;
}

int cp_world_last_timing(cp_world* w, double* ms) {
	if (!w || !ms) return CP_ERR_INVALID;
	for (int i = 0; i < 8; i++) ms[i] = w->last_ms[i];
	return CP_OK;
}

}  // extern "C"
