// Golden-fixture dumper and CPU-baseline timer over the compiled reference
// engine (NathanMacLeod/ConstraintPhysics).  TEST INFRASTRUCTURE ONLY: this
// program is the checker's source of truth, never part of the product path.
//
// It is compiled as ONE translation unit together with the reference's own
// engine sources (the recipe in oracle/Makefile passes the reference root with
// -I).  Member access is opened with `#define private public` so that the
// per-step state (contact cache, constraint graph, island lists) can be read
// without restating any of the reference's orchestration: the only hook into a
// step is a collision action registered through the public API
// (PhysicsEngine::registerCollisionAction, PhysicsEngine.h:110), which the
// reference invokes after its narrowphase and before its island build
// (PhysicsEngine.cpp:340-358 then :364).
//
// Output: a directory holding raw little-endian arrays (<name>.bin) and a
// manifest.json describing dtype/shape of each, plus scalar metadata.

#include <algorithm>
#include <atomic>
#include <cassert>
#include <chrono>
#include <cinttypes>
#include <climits>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <functional>
#include <limits>
#include <map>
#include <mutex>
#include <set>
#include <stack>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>
#include <sys/stat.h>

#define private public
#define protected public
#include "Math/src/Vec3.cpp"
#include "Math/src/Mat3.cpp"
#include "Math/src/NMat.cpp"
#include "ConstraintPhysics/src/CFM.cpp"
#include "ConstraintPhysics/src/CollisionDetect.cpp"
#include "ConstraintPhysics/src/ConstraintSolver.cpp"
#include "ConstraintPhysics/src/ConvexPrimitive.cpp"
#include "ConstraintPhysics/src/Geometry.cpp"
#include "ConstraintPhysics/src/HolonomicBlockSolver.cpp"
#include "ConstraintPhysics/src/Octree.cpp"
#include "ConstraintPhysics/src/PersistentConstraint.cpp"
#include "ConstraintPhysics/src/PhysicsEngine.cpp"
#include "ConstraintPhysics/src/RigidBody.cpp"
#undef private
#undef protected

// Out-of-class definition the reference ODR-uses but never defines.
const int phyz::HolonomicSystem::BLOCK_EMPTY;

using phyz::PhysicsEngine;
using phyz::RigidBody;

// ---------------------------------------------------------------------------
// Seeded generator shared bit-for-bit with the Python/C++ scene builders
// (SplitMix64; u = (x >> 11) * 2^-53).
struct SplitMix64 {
	uint64_t s;
	explicit SplitMix64(uint64_t seed) : s(seed) {}
	uint64_t next() {
		uint64_t z = (s += 0x9E3779B97F4A7C15ull);
		z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
		z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
		return z ^ (z >> 31);
	}
	double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
	double uniform(double a, double b) { return a + (b - a) * uniform(); }
};

// ---------------------------------------------------------------------------
struct Out {
	std::string dir;
	std::vector<std::string> entries;
	std::vector<std::string> metas;

	explicit Out(const std::string& d) : dir(d) { mkdir(d.c_str(), 0755); }

	template <class T>
	void put(const std::string& name, const std::vector<T>& v, std::vector<size_t> shape, const char* dtype) {
		std::string path = dir + "/" + name + ".bin";
		FILE* f = fopen(path.c_str(), "wb");
		if (!f) { perror(path.c_str()); exit(2); }
		if (!v.empty()) fwrite(v.data(), sizeof(T), v.size(), f);
		fclose(f);
		size_t n = 1;
		for (size_t s : shape) n *= s;
		if (n != v.size()) { fprintf(stderr, "shape mismatch for %s\n", name.c_str()); exit(2); }
		std::string e = "\"" + name + "\": {\"dtype\": \"" + dtype + "\", \"shape\": [";
		for (size_t i = 0; i < shape.size(); i++) e += (i ? ", " : "") + std::to_string(shape[i]);
		e += "]}";
		entries.push_back(e);
	}
	void f64(const std::string& n, const std::vector<double>& v, std::vector<size_t> s) { put(n, v, s, "f64"); }
	void i64(const std::string& n, const std::vector<int64_t>& v, std::vector<size_t> s) { put(n, v, s, "i64"); }
	void u64(const std::string& n, const std::vector<uint64_t>& v, std::vector<size_t> s) { put(n, v, s, "u64"); }
	void meta(const std::string& k, double v) {
		char b[64];
		snprintf(b, sizeof b, "%.17g", v);
		metas.push_back("\"" + k + "\": " + b);
	}
	void meta_s(const std::string& k, const std::string& v) { metas.push_back("\"" + k + "\": \"" + v + "\""); }
	void finish() {
		FILE* f = fopen((dir + "/manifest.json").c_str(), "w");
		fprintf(f, "{\n\"arrays\": {\n");
		for (size_t i = 0; i < entries.size(); i++) fprintf(f, "  %s%s\n", entries[i].c_str(), i + 1 < entries.size() ? "," : "");
		fprintf(f, "},\n\"meta\": {\n");
		for (size_t i = 0; i < metas.size(); i++) fprintf(f, "  %s%s\n", metas[i].c_str(), i + 1 < metas.size() ? "," : "");
		fprintf(f, "}\n}\n");
		fclose(f);
	}
};

// ---------------------------------------------------------------------------
// Engine configuration pins (public API only): sleeping off, explicit step
// parameters, every body collidable.  See RigidBody.h:122 / :89 — both flags
// are otherwise left uninitialised by the reference constructors.
struct SceneCfg {
	double dt = 1.0 / 120.0;
	mthz::Vec3 gravity = mthz::Vec3(0, -6, 0);
	int vel_iters = 10, pos_iters = 1, hol_iters = 1;
	int substeps = 8;
};

static void configure(PhysicsEngine& E, const SceneCfg& c) {
	E.setStep_time(c.dt);
	E.setGravity(c.gravity);
	E.setPGSIterations(c.vel_iters, c.pos_iters, c.hol_iters);
	E.setSubstepCount(c.substeps);
	E.setSleepingEnabled(false);
}

static void pin_body(RigidBody* r) {
	r->setNoCollision(false);
	r->setSleepDisabled(true);
}

static phyz::Material c0_material() {
	phyz::Material m = phyz::Material::default_material();
	m.density = 1.0;
	m.restitution = 0.0;
	m.kinetic_friction_coeff = 0.5;
	m.static_friction_coeff = 0.5;
	return m;
}

// Ground: FIXED box((-100,-1,-100), 200, 1, 200) -> top face at y = 0.
// Cubes: n^3 unit cubes, centre (1.1 i, 0.6 + 1.1 j, 1.1 k) + U[-j,j]^3 and yaw
// U[0, 2 pi) about +y, drawn in loop order j (y), i (x), k (z), draws (dx, dy, dz, yaw).
static void build_pile(PhysicsEngine& E, int nx, int ny, int nz, uint64_t seed, double spacing, double jitter) {
	phyz::Material m = c0_material();
	RigidBody* g = E.createRigidBody(phyz::ConvexUnionGeometry::box(mthz::Vec3(-100, -1, -100), 200, 1, 200, m), RigidBody::FIXED);
	pin_body(g);
	SplitMix64 rng(seed);
	for (int j = 0; j < ny; j++)
		for (int i = 0; i < nx; i++)
			for (int k = 0; k < nz; k++) {
				double cx = spacing * i + rng.uniform(-jitter, jitter);
				double cy = 0.6 + spacing * j + rng.uniform(-jitter, jitter);
				double cz = spacing * k + rng.uniform(-jitter, jitter);
				double yaw = rng.uniform(0.0, 2.0 * 3.14159265358979323846);
				RigidBody* r = E.createRigidBody(phyz::ConvexUnionGeometry::box(mthz::Vec3(cx - 0.5, cy - 0.5, cz - 0.5), 1, 1, 1, m));
				r->setOrientation(mthz::Quaternion(yaw, mthz::Vec3(0, 1, 0)));
				pin_body(r);
			}
}

// Mixed pile: like build_pile but every other body is a sphere of radius
// U[0.3, 0.5] (draw after the yaw), so sphere-sphere, sphere-box, box-box and
// body-ground pairs all occur.
// kinds = 3 cycles box, sphere, cylinder: cylinders have r U[0.2, 0.4], h
// U[0.5, 1.0] (draws after the radius) and a random tilt: rotation by U[0, pi)
// about the horizontal axis (cos yaw, 0, sin yaw), so caps, rims and barrels
// all touch.
static void build_mixed(PhysicsEngine& E, int nx, int ny, int nz, uint64_t seed, double spacing, double jitter, int kinds = 2) {
	phyz::Material m = c0_material();
	RigidBody* g = E.createRigidBody(phyz::ConvexUnionGeometry::box(mthz::Vec3(-100, -1, -100), 200, 1, 200, m), RigidBody::FIXED);
	pin_body(g);
	SplitMix64 rng(seed);
	int idx = 0;
	for (int j = 0; j < ny; j++)
		for (int i = 0; i < nx; i++)
			for (int k = 0; k < nz; k++) {
				double cx = spacing * i + rng.uniform(-jitter, jitter);
				double cy = 0.6 + spacing * j + rng.uniform(-jitter, jitter);
				double cz = spacing * k + rng.uniform(-jitter, jitter);
				double yaw = rng.uniform(0.0, 2.0 * 3.14159265358979323846);
				double r = rng.uniform(0.3, 0.5);
				RigidBody* b;
				int kind = idx++ % kinds;
				if (kind == 0) {
					b = E.createRigidBody(phyz::ConvexUnionGeometry::box(mthz::Vec3(cx - 0.5, cy - 0.5, cz - 0.5), 1, 1, 1, m));
					b->setOrientation(mthz::Quaternion(yaw, mthz::Vec3(0, 1, 0)));
				} else if (kind == 1) {
					b = E.createRigidBody(phyz::ConvexUnionGeometry::sphere(mthz::Vec3(cx, cy, cz), r, m));
				} else {
					double cr = 0.2 + (r - 0.3);  // U[0.2, 0.4]
					double h = rng.uniform(0.5, 1.0);
					double tilt = rng.uniform(0.0, 3.14159265358979323846);
					b = E.createRigidBody(phyz::ConvexUnionGeometry::cylinder(mthz::Vec3(cx, cy - h / 2.0, cz), cr, h, m));
					b->setOrientation(mthz::Quaternion(tilt, mthz::Vec3(std::cos(yaw), 0, std::sin(yaw))));
				}
				pin_body(b);
			}
}

// Chains of cylinder links lying along +x, consecutive links joined end to end
// by ball-socket joints, every hinge_every-th joint a hinge about +z (default
// correction stiffnesses, PhysicsEngine.h:15-16).  Links: r 0.15, h 0.8, gap
// 0.1; chain c at z = 1.2 c, height U[0.4, 1.2], x offset U[-0.5, 0.5]; every link
// gets v ~ U[-0.5, 0.5]^3 so the chain bends as it falls onto the ground box.
// Joint kinds as the dumps number them: 1 distance, 3 ball-socket, 4 sliding
// hinge, 5 hinge, 6 weld, 7 slider, 8 cone limit, 9 twist limit, 10 motor,
// 11 piston drive, 12 piston slide limit (a piston is recorded as its slide
// limit, then its drive: the order Piston::addSolverConstraints adds them),
// 13 spring.
struct JointRec {
	int type;
	RigidBody *b1, *b2;
	phyz::PersistentConstraint* pc;
	phyz::Constraint* solver;
	mthz::Vec3 aux = mthz::Vec3();  // a motor: b2's rotation axis of its base hinge (the motor keeps only what it yields);
	                                // a spring: x = the resting length as given (-1: the current distance)
	double ccfm = std::numeric_limits<double>::quiet_NaN();  // setConstraintUseCustomCFM value, NaN: the global CFM
};
static std::vector<JointRec> g_joints;
static RigidBody* g_mesh = nullptr;  // the terrain body of the terrain scene

static void build_chains(PhysicsEngine& E, int n_chains, int n_links, uint64_t seed, int hinge_every) {
	phyz::Material m = c0_material();
	RigidBody* g = E.createRigidBody(phyz::ConvexUnionGeometry::box(mthz::Vec3(-100, -1, -100), 200, 1, 200, m), RigidBody::FIXED);
	pin_body(g);
	SplitMix64 rng(seed);
	const double r = 0.15, h = 0.8, gap = 0.1, PI_ = 3.14159265358979323846;
	mthz::Quaternion q(-PI_ / 2.0, mthz::Vec3(0, 0, 1));  // link axis +y -> +x
	for (int c = 0; c < n_chains; c++) {
		double y0 = rng.uniform(0.4, 1.2), x0 = rng.uniform(-0.5, 0.5), z0 = 1.2 * c;
		RigidBody* prev = nullptr;
		mthz::Vec3 prev_c;
		for (int i = 0; i < n_links; i++) {
			mthz::Vec3 cc(x0 + i * (h + gap), y0, z0);
			RigidBody* b = E.createRigidBody(phyz::ConvexUnionGeometry::cylinder(cc - mthz::Vec3(0, h / 2.0, 0), r, h, m));
			b->setOrientation(q);
			b->setVel(mthz::Vec3(rng.uniform(-0.5, 0.5), rng.uniform(-0.5, 0.5), rng.uniform(-0.5, 0.5)));
			pin_body(b);
			if (prev) {
				mthz::Vec3 a1 = prev_c + mthz::Vec3(0, h / 2.0 + gap / 2.0, 0);
				mthz::Vec3 a2 = cc - mthz::Vec3(0, h / 2.0 + gap / 2.0, 0);
				bool hinge = hinge_every > 0 && i % hinge_every == 0;
				phyz::ConstraintID id = hinge ? E.addHingeConstraint(prev, b, a1, a2, mthz::Vec3(0, 0, 1), mthz::Vec3(0, 0, 1))
				                              : E.addBallSocketConstraint(prev, b, a1, a2);
				auto* e = E.constraint_map[id.getID()];
				phyz::PersistentConstraint* pc = nullptr;
				for (auto* x : e->constraints)
					if (x->getId() == id) pc = x;
				phyz::Constraint* sc = hinge ? (phyz::Constraint*)&static_cast<phyz::Hinge*>(pc)->constraint
				                             : (phyz::Constraint*)&static_cast<phyz::BallSocket*>(pc)->constraint;
				g_joints.push_back({hinge ? 5 : 3, prev, b, pc, sc});
			}
			prev = b;
			prev_c = cc;
		}
	}
}

// Joint solver state: impulse (6, zero padded), the basis u and w of hinges,
// sliders and sliding hinges (zeros for the others).
template <class C>
static void put_impulse(std::vector<double>& s, const C& c, int n) {
	for (int k = 0; k < 6; k++) s.push_back(k < n ? c.impulse.v[k] : 0.0);
}
template <class C>
static void put_joint(std::vector<double>& s, const C& c, int n, bool) {
	put_impulse(s, c, n);
	s.insert(s.end(), {0, 0, 0, 0, 0, 0});
}
template <class C>
static void put_joint_basis(std::vector<double>& s, const C& c, int n) {
	put_impulse(s, c, n);
	s.insert(s.end(), {c.u.x, c.u.y, c.u.z, c.w.x, c.w.y, c.w.z});
}
static void joint_state(std::vector<double>& s) {
	for (auto& j : g_joints) {
		switch (j.type) {
		case 1: put_joint(s, static_cast<phyz::Distance*>(j.pc)->constraint, 1, false); break;
		case 3: put_joint(s, static_cast<phyz::BallSocket*>(j.pc)->constraint, 3, false); break;
		case 4: put_joint_basis(s, static_cast<phyz::SlidingHinge*>(j.pc)->constraint, 4); break;
		case 5: put_joint_basis(s, static_cast<phyz::Hinge*>(j.pc)->constraint, 5); break;
		case 6: put_joint(s, static_cast<phyz::Weld*>(j.pc)->constraint, 6, false); break;
		case 8: {  // impulse, then is_inactive
			auto* c = static_cast<phyz::ConeRotationLimit*>(j.pc);
			put_impulse(s, c->constraint, 1);
			s.insert(s.end(), {c->is_inactive ? 1.0 : 0.0, 0, 0, 0, 0, 0});
			break;
		}
		case 9: {  // impulse, then is_inactive and the tracked current angle
			auto* t = static_cast<phyz::TwistLimit*>(j.pc);
			put_impulse(s, t->constraint, 1);
			s.insert(s.end(), {t->is_inactive ? 1.0 : 0.0, t->current_angle, 0, 0, 0, 0});
			break;
		}
		case 10: {  // impulse, then motor_angular_position and prev_velocity
			auto* m = static_cast<phyz::Motor*>(j.pc);
			put_impulse(s, m->motor_constraint, 1);
			s.insert(s.end(), {m->motor_angular_position, m->prev_velocity, 0, 0, 0, 0});
			break;
		}
		case 11: {  // impulse, then piston_position and prev_velocity
			auto* p = static_cast<phyz::Piston*>(j.pc);
			put_impulse(s, p->piston_constraint, 1);
			s.insert(s.end(), {p->piston_position, p->prev_velocity, 0, 0, 0, 0});
			break;
		}
		case 12: {  // impulse, then !slide_limit_exceeded and piston_position
			auto* p = static_cast<phyz::Piston*>(j.pc);
			put_impulse(s, p->slide_limit, 1);
			s.insert(s.end(), {p->slide_limit_exceeded ? 0.0 : 1.0, p->piston_position, 0, 0, 0, 0});
			break;
		}
		case 13: {  // no solver object: the resting length
			auto* sp = static_cast<phyz::Spring*>(j.pc);
			s.insert(s.end(), {0, 0, 0, 0, 0, 0, sp->resting_length, 0, 0, 0, 0, 0});
			break;
		}
		default: put_joint_basis(s, static_cast<phyz::Slider*>(j.pc)->constraint, 5); break;
		}
	}
}

static int joint_index(const phyz::Constraint* c) {
	for (size_t i = 0; i < g_joints.size(); i++)
		if (g_joints[i].solver == c) return (int)i;
	return -1;
}

// Holonomic systems at this moment: per system the constraint order of
// revaluateSystem (member edges), the elimination order, the block location
// table, vector offsets and the factor buffer (PhysicsEngine.cpp:1785-1798,
// HolonomicBlockSolver.cpp:11-132, :459-543).
struct HsysDump {
	std::vector<int64_t> off{0}, og, order, table_off{0}, table, vloc;
	std::vector<int64_t> buf_off{0};
	std::vector<double> buf;
};
static void dump_hsys(PhysicsEngine& E, HsysDump& d) {
	std::set<PhysicsEngine::HolonomicSystemNodes*> seen;
	std::vector<PhysicsEngine::HolonomicSystemNodes*> list;
	std::vector<unsigned> ids;
	for (auto& kv : E.constraint_graph_nodes) ids.push_back(kv.first);
	std::sort(ids.begin(), ids.end());
	for (unsigned id : ids)
		for (auto* e : E.constraint_graph_nodes[id]->constraints)
			if (e->h && seen.insert(e->h).second) list.push_back(e->h);
	for (auto* h : list) {
		for (auto* e : h->member_edges)
			for (auto* pc : e->constraints)
				if (pc->isHolonomicConstraint()) {
					std::vector<phyz::Constraint*> v;
					pc->addSolverConstraints(&v);
					d.og.push_back(joint_index(v[0]));
				}
		for (auto* c : h->system.constraints) d.order.push_back(joint_index(c));
		d.off.push_back((int64_t)d.order.size());
		for (int t : h->system.block_location_table) d.table.push_back(t);
		d.table_off.push_back((int64_t)d.table.size());
		for (int v : h->system.vector_location_lookup) d.vloc.push_back(v);
		for (int k = 0; k < h->system.buffer_capacity; k++) d.buf.push_back(h->system.buffer[k]);
		d.buf_off.push_back((int64_t)d.buf.size());
	}
}

// ---------------------------------------------------------------------------
// Snapshot helpers.

static std::map<int, std::pair<int, int>> prim_owner(PhysicsEngine& E) {
	std::map<int, std::pair<int, int>> own;
	for (RigidBody* b : E.bodies)
		for (int i = 0; i < (int)b->geometry.size(); i++) own[b->geometry[i].getID()] = {(int)b->getID(), i};
	return own;
}

// cID packs two global primitive ids (CollisionDetect.cpp:827-829); remap each
// half to body_id * 16 + primitive index so that it is stable across runs.
// The half that belongs to a static mesh is a triangle id (CollisionDetect.cpp:
// 2178-2180) and stays as it is: raw_half 0 (low), 1 (high) or -1 (none).
static uint64_t remap_cid(uint64_t cid, const std::map<int, std::pair<int, int>>& own, int raw_half = -1) {
	int a = (int)(cid & 0xFFFFFFFFull), b = (int)(cid >> 32);
	auto fa = own.find(a), fb = own.find(b);
	uint64_t ra = fa == own.end() ? 0xFFFFFFFFull : (uint64_t)(fa->second.first * 16 + fa->second.second);
	uint64_t rb = fb == own.end() ? 0xFFFFFFFFull : (uint64_t)(fb->second.first * 16 + fb->second.second);
	if (raw_half == 0) ra = cid & 0xFFFFFFFFull;
	if (raw_half == 1) rb = cid >> 32;
	return ra | (rb << 32);
}
static int mesh_half(const RigidBody* b1, const RigidBody* b2) {
	if (b1->geometry_type == RigidBody::STATIC_MESH) return 0;
	if (b2->geometry_type == RigidBody::STATIC_MESH) return 1;
	return -1;
}

static void body_state(PhysicsEngine& E, std::vector<double>& s) {
	for (RigidBody* b : E.bodies) {
		mthz::Vec3 c = b->com;
		mthz::Quaternion q = b->orientation;
		s.insert(s.end(), {c.x, c.y, c.z, q.r, q.i, q.j, q.k, b->vel.x, b->vel.y, b->vel.z, b->ang_vel.x, b->ang_vel.y, b->ang_vel.z});
	}
}

static void body_inertia(PhysicsEngine& E, std::vector<double>& s) {
	for (RigidBody* b : E.bodies) {
		for (int r = 0; r < 3; r++) for (int c = 0; c < 3; c++) s.push_back(b->tensor.v[r][c]);
		for (int r = 0; r < 3; r++) for (int c = 0; c < 3; c++) s.push_back(b->invTensor.v[r][c]);
	}
}

struct CacheDump {
	std::vector<int64_t> ints;   // b1, b2, memory_life, is_live, static_ready, pos_in_edge
	std::vector<uint64_t> magic; // cid(remapped), bid
	std::vector<double> f;       // local_b1 3, local_b2 3, normal_local_b1 3, pen, bounce, mu, imp_n, imp_t 2, u 3, w 3
	std::map<const phyz::Constraint*, std::pair<int, int>> rowmap; // solver row -> (cache index, kind)
	int n = 0;
};
static const int CACHE_I = 6, CACHE_F = 21;

static void dump_cache(PhysicsEngine& E, CacheDump& d) {
	auto own = prim_owner(E);
	std::vector<unsigned> ids;
	for (auto& kv : E.constraint_graph_nodes) ids.push_back(kv.first);
	std::sort(ids.begin(), ids.end());
	for (unsigned id : ids) {
		auto* n = E.constraint_graph_nodes[id];
		for (auto* e : n->constraints) {
			auto* other = e->other(n);
			if (other->b->getID() < id) continue;
			int pos = 0;
			for (phyz::PersistentConstraint* pc : e->constraints) {
				if (pc->getId().getType() != phyz::ConstraintID::CONTACT) { pos++; continue; }
				auto* c = static_cast<phyz::Contact*>(pc);
				d.ints.insert(d.ints.end(), {(int64_t)c->b1->getID(), (int64_t)c->b2->getID(), c->memory_life, c->is_live_contact ? 1 : 0,
				                             c->friction.static_ready ? 1 : 0, pos});
				d.magic.push_back(remap_cid(c->magic.cID, own, mesh_half(c->b1, c->b2)));
				d.magic.push_back(c->magic.bID);
				d.f.insert(d.f.end(), {c->contact_pos_local_b1.x, c->contact_pos_local_b1.y, c->contact_pos_local_b1.z,
				                       c->contact_pos_local_b2.x, c->contact_pos_local_b2.y, c->contact_pos_local_b2.z,
				                       c->normal_local_b1.x, c->normal_local_b1.y, c->normal_local_b1.z, c->pen_depth, c->bounce,
				                       c->friction_coeff, c->contact.impulse.v[0], c->friction.impulse.v[0], c->friction.impulse.v[1],
				                       c->friction.u.x, c->friction.u.y, c->friction.u.z, c->friction.w.x, c->friction.w.y, c->friction.w.z});
				d.rowmap[&c->contact] = {d.n, 0};
				d.rowmap[&c->friction] = {d.n, 1};
				d.n++;
				pos++;
			}
		}
	}
}

struct ManifoldDump {
	std::vector<int64_t> pair;   // b1, b2, npoints
	std::vector<double> head;    // normal 3, max_pen
	std::vector<double> pts;     // pos 3, pen
	std::vector<uint64_t> magic; // cid(remapped), bid
};

struct StepDump {
	std::vector<double> pre_state, post_state, pre_inertia, post_inertia;
	CacheDump pre_cache, mid_cache, post_cache;
	ManifoldDump man;
	std::vector<int64_t> pairs;        // broadphase candidate set, (lo, hi) sorted
	std::vector<int64_t> island_rows;  // (cache index, kind) per row
	std::vector<int64_t> island_off;   // row offsets per island
	std::vector<int64_t> node_order;   // constraint_graph_nodes iteration order
	bool mid_done = false;
};

static StepDump* g_cur = nullptr;
static bool g_lite = false;  // omit the mid/post contact caches (large scenes)
static bool g_sleep = false;  // record every body's sleep state after every step
static int g_wake_step = -1, g_wake_body = -1;  // RigidBody::alertWakingAction on that body before that step
static std::function<void(PhysicsEngine&, int)> g_pre_step;  // scene actions before step k (powered-mode changes)
static PhysicsEngine* g_eng = nullptr;

static void on_collision(RigidBody* b1, RigidBody* b2, const std::vector<phyz::Manifold>& ms) {
	StepDump* d = g_cur;
	if (!d) return;
	auto own = prim_owner(*g_eng);
	if (!d->mid_done) {
		d->mid_done = true;
		dump_cache(*g_eng, d->mid_cache);
		for (auto& kv : g_eng->constraint_graph_nodes) d->node_order.push_back(kv.first);
		PhysicsEngine::ActiveConstraintData a = g_eng->sleepOrSolveIslands();
		d->island_off.push_back(0);
		for (auto& isl : a.island_systems) {
			for (phyz::Constraint* c : isl.constraints) {
				auto it = d->mid_cache.rowmap.find(c);
				if (it == d->mid_cache.rowmap.end()) d->island_rows.insert(d->island_rows.end(), {(int64_t)joint_index(c), joint_index(c) >= 0 ? 2 : -1});
				else d->island_rows.insert(d->island_rows.end(), {it->second.first, it->second.second});
			}
			d->island_off.push_back((int64_t)d->island_rows.size() / 2);
		}
	}
	// One call per culled manifold; the list repeats that manifold
	// (PhysicsEngine.cpp:325-333 pushes `man` once per unmerged manifold), so
	// only its first entry is recorded.
	if (!ms.empty()) {
		const phyz::Manifold& m = ms[0];
		d->man.pair.insert(d->man.pair.end(), {(int64_t)b1->getID(), (int64_t)b2->getID(), (int64_t)m.points.size()});
		d->man.head.insert(d->man.head.end(), {m.normal.x, m.normal.y, m.normal.z, m.max_pen_depth});
		for (const phyz::ContactP& p : m.points) {
			d->man.pts.insert(d->man.pts.end(), {p.pos.x, p.pos.y, p.pos.z, p.pen_depth});
			d->man.magic.push_back(remap_cid(p.magicID.cID, own, mesh_half(b1, b2)));
			d->man.magic.push_back(p.magicID.bID);
		}
	}
}

static void broadphase_pairs(PhysicsEngine& E, std::vector<int64_t>& out) {
	std::vector<std::pair<int64_t, int64_t>> v;
	for (auto& p : E.aabb_tree.getAllCollisionCandidates()) {
		int64_t a = p.t1->getID(), b = p.t2->getID();
		if (a > b) std::swap(a, b);
		v.push_back({a, b});
	}
	std::sort(v.begin(), v.end());
	for (auto& p : v) out.insert(out.end(), {p.first, p.second});
}

// Static per-body data: mass properties and, for polyhedra, the body-frame hull
// and its Gauss map exactly as the reference stores them.
static void dump_static(PhysicsEngine& E, Out& o) {
	std::vector<double> mp;
	std::vector<int64_t> bi; // id, movement type, geometry type, n primitives, primitive type
	std::vector<double> pts, gv, gx;
	std::vector<int64_t> poff{0}, foff{0}, fidx{0}, fverts, gred, gint, arcs, aoff{0}, edges, eoff{0}, surfid, ncalc, ndir;
	std::vector<double> mat, sph;  // sphere: body-frame centre xyz, radius
	// cylinder (body frame, as the reference stores it): centre xyz, height axis
	// xyz, radius, height; Gauss vertices [14][3]; cap polygons [8][3] each
	std::vector<double> cyl, cyl_g, cyl_t, cyl_b;
	for (RigidBody* b : E.bodies) {
		mp.push_back(b->mass);
		for (int r = 0; r < 3; r++) for (int c = 0; c < 3; c++) mp.push_back(b->reference_tensor.v[r][c]);
		for (int r = 0; r < 3; r++) for (int c = 0; c < 3; c++) mp.push_back(b->reference_invTensor.v[r][c]);
		int ptype = b->geometry.empty() ? -1 : (int)b->geometry[0].getType();
		bi.insert(bi.end(), {(int64_t)b->getID(), (int64_t)b->movement_type, (int64_t)b->geometry_type, (int64_t)b->geometry.size(), ptype});
		if (!b->geometry.empty()) {
			const phyz::Material& m = b->geometry[0].material;
			mat.insert(mat.end(), {m.density, m.restitution, m.kinetic_friction_coeff, m.static_friction_coeff});
		} else mat.insert(mat.end(), {0, 0, 0, 0});
		if (ptype == phyz::SPHERE && b->geometry.size() == 1) {
			const phyz::Sphere& rs = (const phyz::Sphere&)*b->reference_geometry[0].getGeometry();
			sph.insert(sph.end(), {rs.center.x, rs.center.y, rs.center.z, rs.radius});
		} else sph.insert(sph.end(), {0, 0, 0, 0});
		if (ptype == phyz::CYLINDER && b->geometry.size() == 1) {
			const phyz::Cylinder& rc = (const phyz::Cylinder&)*b->reference_geometry[0].getGeometry();
			cyl.insert(cyl.end(), {rc.center.x, rc.center.y, rc.center.z, rc.height_axis.x, rc.height_axis.y, rc.height_axis.z, rc.radius, rc.height});
			for (auto& v : rc.gauss_verts) cyl_g.insert(cyl_g.end(), {v.x, v.y, v.z});
			for (auto& v : rc.top_face_approximation) cyl_t.insert(cyl_t.end(), {v.x, v.y, v.z});
			for (auto& v : rc.bot_face_approximation) cyl_b.insert(cyl_b.end(), {v.x, v.y, v.z});
			if (rc.gauss_verts.size() != 14 || rc.top_face_approximation.size() != 8) { fprintf(stderr, "unexpected cylinder detail\n"); exit(1); }
		} else {
			cyl.insert(cyl.end(), 8, 0.0);
			cyl_g.insert(cyl_g.end(), 42, 0.0);
			cyl_t.insert(cyl_t.end(), 24, 0.0);
			cyl_b.insert(cyl_b.end(), 24, 0.0);
		}
		if (ptype == phyz::POLYHEDRON && b->geometry.size() == 1) {
			const phyz::Polyhedron& ref = (const phyz::Polyhedron&)*b->reference_geometry[0].getGeometry();
			const phyz::Polyhedron& wrl = (const phyz::Polyhedron&)*b->geometry[0].getGeometry();
			for (auto& p : ref.points) pts.insert(pts.end(), {p.x, p.y, p.z});
			for (auto& s : ref.surfaces) {
				for (uint32_t ix : s.point_indexes) fverts.push_back(ix);
				fidx.push_back((int64_t)fverts.size());
				surfid.push_back(s.getSurfaceID());
				ncalc.push_back(s.normal_calc_index);
				ndir.push_back(s.normalDirection);
			}
			const phyz::GaussMap& g = wrl.gauss_map; // cached extrema live in the world copy
			for (size_t i = 0; i < g.face_verts.size(); i++) {
				const phyz::GaussVert& v = g.face_verts[i];
				const phyz::GaussVert& rv = ref.gauss_map.face_verts[i];
				gv.insert(gv.end(), {rv.v.x, rv.v.y, rv.v.z});
				gred.push_back(v.SAT_redundant ? 1 : 0);
				gint.insert(gint.end(), {v.cached_SAT_query.min_pID, v.cached_SAT_query.max_pID, (int64_t)v.SAT_reference_point_index});
				gx.insert(gx.end(), {v.cached_SAT_query.min_val, v.cached_SAT_query.max_val, v.SAT_reference_point_value});
			}
			for (auto& a : g.arcs) arcs.insert(arcs.end(), {(int64_t)a.v1_indx, (int64_t)a.v2_indx});
			for (auto& e : ref.edges) edges.insert(edges.end(), {e.p1_indx, e.p2_indx});
		}
		poff.push_back((int64_t)pts.size() / 3);
		foff.push_back((int64_t)fidx.size() - 1);
		aoff.push_back((int64_t)arcs.size() / 2);
		eoff.push_back((int64_t)edges.size() / 2);
	}
	size_t N = E.bodies.size();
	o.f64("mass_props", mp, {N, 19});
	o.i64("body_info", bi, {N, 5});
	o.f64("material", mat, {N, 4});
	o.f64("sphere", sph, {N, 4});
	o.f64("cylinder", cyl, {N, 8});
	o.f64("cyl_gauss", cyl_g, {N, 14, 3});
	o.f64("cyl_top", cyl_t, {N, 8, 3});
	o.f64("cyl_bot", cyl_b, {N, 8, 3});
	o.f64("hull_points", pts, {pts.size() / 3, 3});
	o.i64("hull_point_off", poff, {N + 1});
	o.i64("hull_face_off", foff, {N + 1});
	o.i64("hull_face_idx", fidx, {fidx.size()});
	o.i64("hull_face_verts", fverts, {fverts.size()});
	o.i64("hull_surface_id", surfid, {surfid.size()});
	o.i64("hull_normal_calc", ncalc, {ncalc.size()});
	o.i64("hull_normal_dir", ndir, {ndir.size()});
	o.f64("gauss_normal", gv, {gv.size() / 3, 3});
	o.i64("gauss_redundant", gred, {gred.size()});
	o.i64("gauss_ints", gint, {gint.size() / 3, 3});
	o.f64("gauss_vals", gx, {gx.size() / 3, 3});
	o.i64("gauss_arcs", arcs, {arcs.size() / 2, 2});
	o.i64("gauss_arc_off", aoff, {N + 1});
	o.i64("hull_edges", edges, {edges.size() / 2, 2});
	o.i64("hull_edge_off", eoff, {N + 1});
	std::vector<int64_t> ji;
	std::vector<double> jg;
	// geometry per joint: tracked points (body frames), two body-frame axes
	// (hinge axes, slide axes; a distance joint's target distance in the first
	// slot), position and rotation correction strengths
	auto geo = [&](const JointRec& j, int k1, int k2, mthz::Vec3 a1, mthz::Vec3 a2, double ps, double rs) {
		mthz::Vec3 t1 = j.b1->track_p[k1], t2 = j.b2->track_p[k2];
		jg.insert(jg.end(), {t1.x, t1.y, t1.z, t2.x, t2.y, t2.z, a1.x, a1.y, a1.z, a2.x, a2.y, a2.z, ps, rs});
	};
	for (auto& j : g_joints) {
		ji.insert(ji.end(), {(int64_t)j.type, (int64_t)j.b1->getID(), (int64_t)j.b2->getID()});
		switch (j.type) {
		case 1: {
			auto* d = static_cast<phyz::Distance*>(j.pc);
			geo(j, d->b1_point_key, d->b2_point_key, mthz::Vec3(d->target_distance, 0, 0), mthz::Vec3(), d->pos_correct_hardness, 0);
			break;
		}
		case 3: {
			auto* b = static_cast<phyz::BallSocket*>(j.pc);
			geo(j, b->b1_point_key, b->b2_point_key, mthz::Vec3(), mthz::Vec3(), b->pos_correct_hardness, 0);
			break;
		}
		case 4: {
			auto* h = static_cast<phyz::SlidingHinge*>(j.pc);
			geo(j, h->b1_point_key, h->b2_point_key, h->b1_slide_axis_body_space, h->b2_slide_axis_body_space, h->pos_correct_hardness,
			    h->rot_correct_hardness);
			break;
		}
		case 5: {
			auto* h = static_cast<phyz::Hinge*>(j.pc);
			geo(j, h->b1_point_key, h->b2_point_key, h->b1_rot_axis_body_space, h->b2_rot_axis_body_space, h->pos_correct_hardness,
			    h->rot_correct_hardness);
			break;
		}
		case 8: {  // tracked-point slots: the cone's max angle
			auto* c = static_cast<phyz::ConeRotationLimit*>(j.pc);
			mthz::Vec3 a1 = c->b1_allignment_axis_local, a2 = c->b2_allignment_axis_local;
			jg.insert(jg.end(), {c->max_rotation_angle, 0, 0, 0, 0, 0, a1.x, a1.y, a1.z, a2.x, a2.y, a2.z, 0, c->rot_correct_hardness});
			break;
		}
		case 9: {  // tracked-point slots: the twist's min and max angles
			auto* t = static_cast<phyz::TwistLimit*>(j.pc);
			mthz::Vec3 a1 = t->b1_twist_axis_local, a2 = t->b2_twist_axis_local;
			jg.insert(jg.end(), {t->min_angle, t->max_angle, 0, 0, 0, 0, a1.x, a1.y, a1.z, a2.x, a2.y, a2.z, 0, t->rot_correct_hardness});
			break;
		}
		case 10: {  // tracked-point slots: the motor's min and max angles
			auto* m = static_cast<phyz::Motor*>(j.pc);
			mthz::Vec3 a1 = m->b1_rot_axis_body_space, a2 = j.aux;
			jg.insert(jg.end(), {m->min_motor_position, m->max_motor_position, 0, 0, 0, 0, a1.x, a1.y, a1.z, a2.x, a2.y, a2.z, 0,
			                     m->rot_correct_hardness});
			break;
		}
		case 11:
		case 12: {  // axis slots: b1's slide axis, then the slide limits
			auto* p = static_cast<phyz::Piston*>(j.pc);
			mthz::Vec3 t1 = j.b1->track_p[p->b1_point_key], t2 = j.b2->track_p[p->b2_point_key];
			mthz::Vec3 a1 = p->b1_slide_axis_body_space;
			jg.insert(jg.end(), {t1.x, t1.y, t1.z, t2.x, t2.y, t2.z, a1.x, a1.y, a1.z, p->min_slide_limit, p->max_slide_limit, 0,
			                     p->pos_correct_hardness, 0});
			break;
		}
		case 13: {  // axis slots: stiffness, damping, resting length as given
			auto* sp = static_cast<phyz::Spring*>(j.pc);
			geo(j, sp->b1_point_key, sp->b2_point_key, mthz::Vec3(sp->stiffness, sp->damping, j.aux.x), mthz::Vec3(), 0, 0);
			break;
		}
		case 6: {
			auto* w = static_cast<phyz::Weld*>(j.pc);
			geo(j, w->b1_point_key, w->b2_point_key, mthz::Vec3(), mthz::Vec3(), w->pos_correct_hardness, w->rot_correct_hardness);
			break;
		}
		default: {
			auto* h = static_cast<phyz::Slider*>(j.pc);
			geo(j, h->b1_point_key, h->b2_point_key, h->b1_slide_axis_body_space, h->b2_slide_axis_body_space, h->pos_correct_hardness,
			    h->rot_correct_hardness);
			break;
		}
		}
	}
	o.i64("joint_info", ji, {g_joints.size(), 3});
	o.f64("joint_geo", jg, {g_joints.size(), 14});
	std::vector<double> jc;
	for (auto& j : g_joints) jc.push_back(j.ccfm);
	o.f64("joint_cfm", jc, {g_joints.size()});
}

static void put_cache(Out& o, const std::string& p, const CacheDump& c) {
	o.i64(p + "_i", c.ints, {(size_t)c.n, (size_t)CACHE_I});
	o.u64(p + "_m", c.magic, {(size_t)c.n, 2});
	o.f64(p + "_f", c.f, {(size_t)c.n, (size_t)CACHE_F});
}

// Run `steps` steps; record full dumps at the listed step indices.
static void run_trace(PhysicsEngine& E, const SceneCfg& cfg, int steps, const std::set<int>& record, Out& o) {
	size_t N = E.bodies.size();
	dump_static(E, o);
	{
		std::vector<double> s;
		body_state(E, s);
		o.f64("initial_state", s, {N, 13});
	}
	o.meta("n_bodies", (double)N);
	o.meta("dt", cfg.dt);
	o.meta("gravity_y", cfg.gravity.y);
	o.meta("vel_iters", cfg.vel_iters);
	o.meta("pos_iters", cfg.pos_iters);
	o.meta("substeps", cfg.substeps);
	o.meta("hol_iters", cfg.hol_iters);
	o.meta("steps", steps);
	std::vector<int64_t> rec_list;
	std::vector<double> sleep_trace;
	g_eng = &E;
	E.registerCollisionAction(phyz::CollisionTarget::all(), phyz::CollisionTarget::all(), on_collision);
	for (int k = 0; k < steps; k++) {
		bool rec = record.count(k) > 0;
		StepDump d;
		std::vector<double> jpre;
		HsysDump hs;
		std::vector<int64_t> c_off{0}, c_list;
		std::vector<double> c_box;
		if (rec && g_mesh) {  // mesh candidates for each body's box before the step
			for (RigidBody* b : E.bodies) {
				if (b != g_mesh)
					for (unsigned t : g_mesh->mesh.getAABBTree().getCollisionCandidatesWith(b->aabb)) c_list.push_back(t);
				c_off.push_back((int64_t)c_list.size());
				c_box.insert(c_box.end(), {b->aabb.min.x, b->aabb.min.y, b->aabb.min.z, b->aabb.max.x, b->aabb.max.y, b->aabb.max.z});
			}
		}
		if (rec) {
			joint_state(jpre);
			dump_hsys(E, hs);
			body_state(E, d.pre_state);
			body_inertia(E, d.pre_inertia);
			dump_cache(E, d.pre_cache);
			g_cur = &d;
		}
		if (k == g_wake_step) E.bodies[g_wake_body]->alertWakingAction();
		if (g_pre_step) g_pre_step(E, k);
		E.timeStep();
		g_cur = nullptr;
		if (g_sleep)
			for (RigidBody* b : E.bodies)
				sleep_trace.insert(sleep_trace.end(), {b->asleep ? 1.0 : 0.0, b->sleep_ready_counter, (double)b->non_sleepy_tick_count,
				                                       (double)b->history.size()});
		if (!rec) continue;
		rec_list.push_back(k);
		broadphase_pairs(E, d.pairs);
		body_state(E, d.post_state);
		body_inertia(E, d.post_inertia);
		dump_cache(E, d.post_cache);
		std::string p = "s" + std::to_string(k);
		o.f64(p + "_pre", d.pre_state, {N, 13});
		o.f64(p + "_post", d.post_state, {N, 13});
		o.f64(p + "_pre_inertia", d.pre_inertia, {N, 18});
		o.f64(p + "_post_inertia", d.post_inertia, {N, 18});
		put_cache(o, p + "_cache_pre", d.pre_cache);
		if (!g_lite) {
			put_cache(o, p + "_cache_mid", d.mid_cache);
			put_cache(o, p + "_cache_post", d.post_cache);
		}
		o.i64(p + "_pairs", d.pairs, {d.pairs.size() / 2, 2});
		o.i64(p + "_man_pair", d.man.pair, {d.man.pair.size() / 3, 3});
		o.f64(p + "_man_head", d.man.head, {d.man.head.size() / 4, 4});
		o.f64(p + "_man_pts", d.man.pts, {d.man.pts.size() / 4, 4});
		o.u64(p + "_man_magic", d.man.magic, {d.man.magic.size() / 2, 2});
		o.i64(p + "_island_rows", d.island_rows, {d.island_rows.size() / 2, 2});
		if (d.island_off.empty()) d.island_off.push_back(0);
		o.i64(p + "_island_off", d.island_off, {d.island_off.size()});
		o.i64(p + "_node_order", d.node_order, {d.node_order.size()});
		if (g_mesh) {
			o.i64(p + "_mesh_cand_off", c_off, {c_off.size()});
			o.i64(p + "_mesh_cand", c_list, {c_list.size()});
			o.f64(p + "_mesh_cand_box", c_box, {N, 6});
		}
		if (!g_joints.empty()) {
			std::vector<double> jpost;
			joint_state(jpost);
			o.f64(p + "_joint_pre", jpre, {g_joints.size(), 12});
			o.f64(p + "_joint_post", jpost, {g_joints.size(), 12});
			o.i64(p + "_hsys_off", hs.off, {hs.off.size()});
			o.i64(p + "_hsys_og", hs.og, {hs.og.size()});
			o.i64(p + "_hsys_order", hs.order, {hs.order.size()});
			o.i64(p + "_hsys_table_off", hs.table_off, {hs.table_off.size()});
			o.i64(p + "_hsys_table", hs.table, {hs.table.size()});
			o.i64(p + "_hsys_vloc", hs.vloc, {hs.vloc.size()});
			o.i64(p + "_hsys_buf_off", hs.buf_off, {hs.buf_off.size()});
			o.f64(p + "_hsys_buf", hs.buf, {hs.buf.size()});
		}
	}
	o.i64("recorded_steps", rec_list, {rec_list.size()});
	if (g_sleep) o.f64("sleep_trace", sleep_trace, {(size_t)steps, N, 4});
	{
		std::vector<double> s;
		body_state(E, s);
		o.f64("final_state", s, {N, 13});
	}
}

// ---------------------------------------------------------------------------
static double arg_d(int argc, char** argv, const char* key, double def) {
	for (int i = 1; i + 1 < argc; i++) if (!strcmp(argv[i], key)) return atof(argv[i + 1]);
	return def;
}
static std::string arg_s(int argc, char** argv, const char* key, const char* def) {
	for (int i = 1; i + 1 < argc; i++) if (!strcmp(argv[i], key)) return argv[i + 1];
	return def;
}
static std::set<int> arg_steps(int argc, char** argv, const char* key) {
	std::set<int> s;
	std::string v = arg_s(argc, argv, key, "");
	size_t pos = 0;
	while (pos < v.size()) {
		size_t c = v.find(',', pos);
		if (c == std::string::npos) c = v.size();
		s.insert(atoi(v.substr(pos, c - pos).c_str()));
		pos = c + 1;
	}
	return s;
}

// pile: C0-pattern box pile (n x n x n), trace dump.
static int cmd_pile(int argc, char** argv) {
	int n = (int)arg_d(argc, argv, "--n", 3);
	int ny = (int)arg_d(argc, argv, "--ny", n);
	int steps = (int)arg_d(argc, argv, "--steps", 10);
	uint64_t seed = (uint64_t)arg_d(argc, argv, "--seed", 12345);
	SceneCfg cfg;
	g_lite = arg_d(argc, argv, "--lite", 0) != 0;
	cfg.vel_iters = (int)arg_d(argc, argv, "--vel", 10);
	cfg.substeps = (int)arg_d(argc, argv, "--substeps", 8);
	PhysicsEngine E;
	configure(E, cfg);
	if (arg_d(argc, argv, "--mixed", 0) != 0)
		build_mixed(E, n, ny, n, seed, arg_d(argc, argv, "--spacing", 1.1), arg_d(argc, argv, "--jitter", 0.05),
		            arg_d(argc, argv, "--mixed", 0) >= 2 ? 3 : 2);
	else
		build_pile(E, n, ny, n, seed, arg_d(argc, argv, "--spacing", 1.1), arg_d(argc, argv, "--jitter", 0.05));
	Out o(arg_s(argc, argv, "--out", "out"));
	o.meta("seed", (double)seed);
	o.meta("nx", n);
	o.meta("ny", ny);
	run_trace(E, cfg, steps, arg_steps(argc, argv, "--record"), o);
	o.finish();
	return 0;
}

// chain: jointed cylinder chains on a ground box, trace dump.
static int cmd_chain(int argc, char** argv) {
	int chains = (int)arg_d(argc, argv, "--chains", 3);
	int links = (int)arg_d(argc, argv, "--links", 6);
	int steps = (int)arg_d(argc, argv, "--steps", 10);
	uint64_t seed = (uint64_t)arg_d(argc, argv, "--seed", 4242);
	SceneCfg cfg;
	cfg.hol_iters = (int)arg_d(argc, argv, "--hol", 1);
	PhysicsEngine E;
	configure(E, cfg);
	build_chains(E, chains, links, seed, (int)arg_d(argc, argv, "--hinge-every", 3));
	Out o(arg_s(argc, argv, "--out", "out"));
	o.meta("seed", (double)seed);
	run_trace(E, cfg, steps, arg_steps(argc, argv, "--record"), o);
	o.finish();
	return 0;
}

// Heightfield terrain: generateGridMeshInput(grid, grid, cell) (Geometry.cpp:
// 589-611) centred on the origin, point heights amp * value noise (a lattice
// of U[-1, 1] values every 4 cells, bilinear with smoothstep weights).  The
// mesh body is FIXED with its centre of mass pinned at the origin (the
// reference leaves it unset for meshes, RigidBody.cpp:40-58).  Above it a
// mixed pile (box, sphere, tilted cylinder in turn; nx x ny x nz, spacing 1.1)
// and `chains` chains of `links` cylinder links as in build_chains.
static phyz::MeshInput g_mesh_input;
static void build_terrain(PhysicsEngine& E, int grid, double cell, double amp, int nx, int ny, int nz, int n_chains, int n_links,
                          uint64_t seed) {
	phyz::Material m = c0_material();
	double half = 0.5 * grid * cell;
	phyz::MeshInput in = phyz::generateGridMeshInput(grid, grid, cell, mthz::Vec3(-half, 0, -half), m);
	SplitMix64 rng(seed);
	int L = grid / 4 + 2;
	std::vector<double> lat(L * L);
	for (double& v : lat) v = rng.uniform(-1.0, 1.0);
	auto smooth = [](double t) { return t * t * (3 - 2 * t); };
	for (int j = 0; j <= grid; j++)
		for (int i = 0; i <= grid; i++) {
			double gx = i / 4.0, gz = j / 4.0;
			int ix = (int)gx, iz = (int)gz;
			double fx = smooth(gx - ix), fz = smooth(gz - iz);
			double a = lat[ix + L * iz], b = lat[ix + 1 + L * iz], c = lat[ix + L * (iz + 1)], d = lat[ix + 1 + L * (iz + 1)];
			double h = (a * (1 - fx) + b * fx) * (1 - fz) + (c * (1 - fx) + d * fx) * fz;
			in.points[i + (grid + 1) * j].y = amp * h;
		}
	g_mesh_input = in;
	RigidBody* g = E.createRigidBody(phyz::StaticMeshGeometry(in), true);
	g->setCOMType(RigidBody::PHYSICALLY_BASED);
	g->translateSoCOMAtPosition(mthz::Vec3(0, 0, 0));
	pin_body(g);
	g_mesh = g;
	double y0 = amp + 0.6;
	int idx = 0;
	for (int j = 0; j < ny; j++)
		for (int i = 0; i < nx; i++)
			for (int k = 0; k < nz; k++) {
				double cx = -0.5 * 1.1 * (nx - 1) + 1.1 * i + rng.uniform(-0.05, 0.05);
				double cy = y0 + 1.1 * j + rng.uniform(-0.05, 0.05);
				double cz = -0.5 * 1.1 * (nz - 1) + 1.1 * k + rng.uniform(-0.05, 0.05);
				double yaw = rng.uniform(0.0, 2.0 * 3.14159265358979323846);
				double r = rng.uniform(0.3, 0.5);
				RigidBody* b;
				int kind = idx++ % 3;
				if (kind == 0) {
					b = E.createRigidBody(phyz::ConvexUnionGeometry::box(mthz::Vec3(cx - 0.5, cy - 0.5, cz - 0.5), 1, 1, 1, m));
					b->setOrientation(mthz::Quaternion(yaw, mthz::Vec3(0, 1, 0)));
				} else if (kind == 1) {
					b = E.createRigidBody(phyz::ConvexUnionGeometry::sphere(mthz::Vec3(cx, cy, cz), r, m));
				} else {
					double cr = 0.2 + (r - 0.3);
					double h = rng.uniform(0.5, 1.0);
					double tilt = rng.uniform(0.0, 3.14159265358979323846);
					b = E.createRigidBody(phyz::ConvexUnionGeometry::cylinder(mthz::Vec3(cx, cy - h / 2.0, cz), cr, h, m));
					b->setOrientation(mthz::Quaternion(tilt, mthz::Vec3(std::cos(yaw), 0, std::sin(yaw))));
				}
				pin_body(b);
			}
	const double r = 0.15, h = 0.8, gap = 0.1, PI_ = 3.14159265358979323846;
	mthz::Quaternion q(-PI_ / 2.0, mthz::Vec3(0, 0, 1));
	for (int c = 0; c < n_chains; c++) {
		double yc = y0 + 1.1 * ny + rng.uniform(0.0, 0.4), x0 = -0.5 * n_links * (h + gap) + rng.uniform(-0.3, 0.3);
		double z0 = -0.5 * 1.2 * (n_chains - 1) + 1.2 * c;
		RigidBody* prev = nullptr;
		mthz::Vec3 prev_c;
		for (int i = 0; i < n_links; i++) {
			mthz::Vec3 cc(x0 + i * (h + gap), yc, z0);
			RigidBody* b = E.createRigidBody(phyz::ConvexUnionGeometry::cylinder(cc - mthz::Vec3(0, h / 2.0, 0), r, h, m));
			b->setOrientation(q);
			b->setVel(mthz::Vec3(rng.uniform(-0.5, 0.5), rng.uniform(-0.5, 0.5), rng.uniform(-0.5, 0.5)));
			pin_body(b);
			if (prev) {
				mthz::Vec3 a1 = prev_c + mthz::Vec3(0, h / 2.0 + gap / 2.0, 0);
				mthz::Vec3 a2 = cc - mthz::Vec3(0, h / 2.0 + gap / 2.0, 0);
				bool hinge = i % 3 == 0;
				phyz::ConstraintID id = hinge ? E.addHingeConstraint(prev, b, a1, a2, mthz::Vec3(0, 0, 1), mthz::Vec3(0, 0, 1))
				                              : E.addBallSocketConstraint(prev, b, a1, a2);
				auto* e = E.constraint_map[id.getID()];
				phyz::PersistentConstraint* pc = nullptr;
				for (auto* x : e->constraints)
					if (x->getId() == id) pc = x;
				phyz::Constraint* sc = hinge ? (phyz::Constraint*)&static_cast<phyz::Hinge*>(pc)->constraint
				                             : (phyz::Constraint*)&static_cast<phyz::BallSocket*>(pc)->constraint;
				g_joints.push_back({hinge ? 5 : 3, prev, b, pc, sc});
			}
			prev = b;
			prev_c = cc;
		}
	}
}

// The mesh as built by StaticMeshGeometry: input points and triangles, then per
// world triangle the derived data (normal, vertices, edge out directions,
// Gauss region; vertex and edge ids, region size, concave neighbours).
static void dump_mesh(Out& o, const phyz::MeshInput& in) {
	std::vector<double> pts, tmat, tf;
	std::vector<int64_t> tidx, ti;
	for (auto& p : in.points) pts.insert(pts.end(), {p.x, p.y, p.z});
	for (auto& t : in.triangle_indices) {
		tidx.insert(tidx.end(), {(int64_t)t.i1, (int64_t)t.i2, (int64_t)t.i3});
		tmat.insert(tmat.end(), {t.material.restitution, t.material.kinetic_friction_coeff, t.material.static_friction_coeff});
	}
	for (const phyz::StaticMeshFace& f : g_mesh->mesh.getTriangles()) {
		auto put = [&](mthz::Vec3 v) { tf.insert(tf.end(), {v.x, v.y, v.z}); };
		put(f.normal);
		for (int k = 0; k < 3; k++) put(f.vertices[k].p);
		for (int k = 0; k < 3; k++) put(f.edges[k].out_direction);
		for (int k = 0; k < 3; k++) put(k < (int)f.gauss_region.size() ? f.gauss_region[k] : mthz::Vec3(0, 0, 0));
		tf.insert(tf.end(), {f.aabb.min.x, f.aabb.min.y, f.aabb.min.z, f.aabb.max.x, f.aabb.max.y, f.aabb.max.z});
		for (int k = 0; k < 3; k++) ti.push_back(f.vertices[k].id);
		for (int k = 0; k < 3; k++) ti.push_back(f.edges[k].id);
		ti.insert(ti.end(), {(int64_t)f.gauss_region.size(), (int64_t)f.concave_neighbor_count, (int64_t)f.id});
	}
	size_t T = in.triangle_indices.size();
	o.f64("mesh_points", pts, {in.points.size(), 3});
	o.i64("mesh_tris", tidx, {T, 3});
	o.f64("mesh_tri_mat", tmat, {T, 3});
	o.f64("mesh_tri_f", tf, {T, 36});
	o.i64("mesh_tri_i", ti, {T, 9});
}


// Rays cast at the state after the last step through the public query
// (PhysicsEngine::raycastFirstIntersection, PhysicsEngine.cpp:582-609):
// n_rays rays from points 6 m from a random body's centre (raised 3 m),
// aimed at that centre with U[-0.3, 0.3]^3 jitter; every 8th ray straight
// down (0, -1, 0) (the axis-parallel branch of rayIntersectsAABB), every
// 16th in a random direction.  Dumps origins, unit directions, (hit, body id)
// and hit point, surface normal, distance.
static void dump_rays(PhysicsEngine& E, Out& o, int n_rays, uint64_t seed) {
	SplitMix64 rng(seed ^ 0x5A5A5A5Aull);
	std::vector<double> org, dir, res;
	std::vector<int64_t> hit;
	int N = (int)E.bodies.size();
	for (int r = 0; r < n_rays; r++) {
		RigidBody* tb = E.bodies[(size_t)(rng.next() % (uint64_t)N)];
		mthz::Vec3 c = tb->getCOM();
		if (tb->geometry_type == RigidBody::STATIC_MESH) c = mthz::Vec3(rng.uniform(-2, 2), 0, rng.uniform(-2, 2));
		mthz::Vec3 off(rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(-1, 1));
		mthz::Vec3 jit(rng.uniform(-0.3, 0.3), rng.uniform(-0.3, 0.3), rng.uniform(-0.3, 0.3));
		mthz::Vec3 rnd(rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(-1, 1));
		mthz::Vec3 origin = c + off.normalize() * 6.0 + mthz::Vec3(0, 3, 0);
		mthz::Vec3 d = (c + jit - origin).normalize();
		if (r % 8 == 7) d = mthz::Vec3(0, -1, 0);
		if (r % 16 == 15) d = rnd.normalize();
		phyz::RayHitInfo h = E.raycastFirstIntersection(origin, d);
		org.insert(org.end(), {origin.x, origin.y, origin.z});
		dir.insert(dir.end(), {d.x, d.y, d.z});
		hit.insert(hit.end(), {(int64_t)h.did_hit, h.did_hit ? (int64_t)h.hit_object->getID() : 0});
		if (h.did_hit)
			res.insert(res.end(), {h.hit_position.x, h.hit_position.y, h.hit_position.z, h.surface_normal.x, h.surface_normal.y,
			                       h.surface_normal.z, h.hit_distance});
		else
			res.insert(res.end(), 7, 0.0);
	}
	o.f64("ray_origin", org, {(size_t)n_rays, 3});
	o.f64("ray_dir", dir, {(size_t)n_rays, 3});
	o.i64("ray_hit", hit, {(size_t)n_rays, 2});
	o.f64("ray_out", res, {(size_t)n_rays, 7});
}

// terrain: mixed bodies and jointed chains on a heightfield mesh, trace dump.
static int cmd_terrain(int argc, char** argv) {
	int grid = (int)arg_d(argc, argv, "--grid", 12);
	double cell = arg_d(argc, argv, "--cell", 0.5);
	double amp = arg_d(argc, argv, "--amp", 0.3);
	int n = (int)arg_d(argc, argv, "--n", 2);
	int ny = (int)arg_d(argc, argv, "--ny", n);
	int steps = (int)arg_d(argc, argv, "--steps", 10);
	uint64_t seed = (uint64_t)arg_d(argc, argv, "--seed", 777);
	SceneCfg cfg;
	cfg.hol_iters = (int)arg_d(argc, argv, "--hol", 1);
	PhysicsEngine E;
	configure(E, cfg);
	build_terrain(E, grid, cell, amp, n, ny, n, (int)arg_d(argc, argv, "--chains", 2), (int)arg_d(argc, argv, "--links", 4), seed);
	Out o(arg_s(argc, argv, "--out", "out"));
	o.meta("seed", (double)seed);
	o.meta("grid", grid);
	dump_mesh(o, g_mesh_input);
	std::set<int> rec = arg_steps(argc, argv, "--record");
	run_trace(E, cfg, steps, rec, o);
	dump_rays(E, o, (int)arg_d(argc, argv, "--rays", 0), seed);
	o.finish();
	return 0;
}

// joints: the equality joints beyond ball-sockets and hinges.  A chain hangs
// from a FIXED anchor box: anchor -distance- box -weld- box -slider- box
// -sliding hinge- cylinder (one holonomic system of degrees 1, 6, 5, 4); a
// welded pair of boxes and a slider pair (a box sliding along another, axis
// tilted) fall onto the ground box.  Every dynamic body gets a random
// velocity U[-1, 1]^3 and angular velocity U[-1, 1]^3; default material and
// correction strengths (PhysicsEngine.h:15-16).
template <class T>
static void record_joint(PhysicsEngine& E, phyz::ConstraintID id, int type, RigidBody* b1, RigidBody* b2) {
	auto* e = E.constraint_map[id.getID()];
	phyz::PersistentConstraint* pc = nullptr;
	for (auto* x : e->constraints)
		if (x->getId() == id) pc = x;
	g_joints.push_back({type, b1, b2, pc, (phyz::Constraint*)&static_cast<T*>(pc)->constraint});
}
static int cmd_joints(int argc, char** argv) {
	int steps = (int)arg_d(argc, argv, "--steps", 120);
	uint64_t seed = (uint64_t)arg_d(argc, argv, "--seed", 909);
	SceneCfg cfg;
	cfg.hol_iters = (int)arg_d(argc, argv, "--hol", 1);
	PhysicsEngine E;
	configure(E, cfg);
	phyz::Material m = c0_material();
	RigidBody* g = E.createRigidBody(phyz::ConvexUnionGeometry::box(mthz::Vec3(-100, -1, -100), 200, 1, 200, m), RigidBody::FIXED);
	pin_body(g);
	SplitMix64 rng(seed);
	auto kick = [&](RigidBody* b) {
		b->setVel(mthz::Vec3(rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(-1, 1)));
		b->setAngVel(mthz::Vec3(rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(-1, 1)));
		pin_body(b);
	};
	auto box = [&](mthz::Vec3 c, double sx, double sy, double sz) {
		return E.createRigidBody(phyz::ConvexUnionGeometry::box(c - mthz::Vec3(sx / 2, sy / 2, sz / 2), sx, sy, sz, m));
	};
	// the hanging chain (bodies centred on x = 0, z = 0)
	RigidBody* anchor = E.createRigidBody(phyz::ConvexUnionGeometry::box(mthz::Vec3(-0.5, 9.5, -0.5), 1, 1, 1, m), RigidBody::FIXED);
	pin_body(anchor);
	RigidBody* l1 = box(mthz::Vec3(0.3, 8.2, 0), 0.6, 0.6, 0.6);
	RigidBody* l2 = box(mthz::Vec3(0.3, 7.5, 0), 0.6, 0.6, 0.6);
	RigidBody* l3 = box(mthz::Vec3(0.3, 6.8, 0.1), 0.5, 0.5, 0.5);
	RigidBody* l4 = E.createRigidBody(phyz::ConvexUnionGeometry::cylinder(mthz::Vec3(0.3, 5.6, 0.1), 0.25, 0.8, m));
	for (RigidBody* b : {l1, l2, l3, l4}) kick(b);
	record_joint<phyz::Distance>(E, E.addDistanceConstraint(anchor, l1, mthz::Vec3(0, -0.5, 0), mthz::Vec3(-0.2, 0.3, 0)), 1, anchor, l1);
	record_joint<phyz::Weld>(E, E.addWeldConstraint(l1, l2, mthz::Vec3(0, -0.35, 0), mthz::Vec3(0, 0.35, 0)), 6, l1, l2);
	record_joint<phyz::Slider>(E, E.addSliderConstraint(l2, l3, mthz::Vec3(0, -0.3, 0), mthz::Vec3(0, 0.25, 0), mthz::Vec3(0, 1, 0),
	                                                    mthz::Vec3(0, 1, 0)), 7, l2, l3);
	record_joint<phyz::SlidingHinge>(E, E.addSlidingHingeConstraint(l3, l4, mthz::Vec3(0, -0.25, 0), mthz::Vec3(0, 0.4, 0),
	                                                                mthz::Vec3(0, 1, 0), mthz::Vec3(0, 1, 0)), 4, l3, l4);
	// a welded pair and a slider pair on the ground
	RigidBody* w1 = box(mthz::Vec3(4, 1.5, 0), 1, 0.5, 1);
	RigidBody* w2 = box(mthz::Vec3(4.9, 1.8, 0), 0.8, 0.8, 0.8);
	RigidBody* s1 = box(mthz::Vec3(-4, 1.2, 0), 2, 0.4, 1);
	RigidBody* s2 = box(mthz::Vec3(-4.3, 1.7, 0), 0.5, 0.5, 0.5);
	for (RigidBody* b : {w1, w2, s1, s2}) kick(b);
	record_joint<phyz::Weld>(E, E.addWeldConstraint(w1, w2, mthz::Vec3(0.5, 0.1, 0), mthz::Vec3(-0.4, -0.2, 0)), 6, w1, w2);
	record_joint<phyz::Slider>(E, E.addSliderConstraint(s1, s2, mthz::Vec3(0, 0.2, 0), mthz::Vec3(0, -0.25, 0), mthz::Vec3(1, 0.1, 0),
	                                                    mthz::Vec3(1, 0.1, 0)), 7, s1, s2);
	Out o(arg_s(argc, argv, "--out", "out"));
	o.meta("seed", (double)seed);
	run_trace(E, cfg, steps, arg_steps(argc, argv, "--record"), o);
	o.finish();
	return 0;
}

// limits: cone and twist limits.  Two pendulums hang from FIXED anchors:
// a two-link ball-socket pendulum (one holonomic system) with a cone limit
// (0.12 rad about the downward axis) between its links and one (0.15 rad)
// between the anchor and the first link; and a box on a hinge about +y with a twist
// limit [-0.3, 0.4] rad about the same axis.  Each link gets a sideways
// velocity U[1, 2] and an angular velocity about +y U[2, 3] so the limits
// engage and release; default correction strengths.
static int cmd_limits(int argc, char** argv) {
	int steps = (int)arg_d(argc, argv, "--steps", 150);
	uint64_t seed = (uint64_t)arg_d(argc, argv, "--seed", 311);
	SceneCfg cfg;
	cfg.hol_iters = (int)arg_d(argc, argv, "--hol", 1);
	PhysicsEngine E;
	configure(E, cfg);
	phyz::Material m = c0_material();
	RigidBody* g = E.createRigidBody(phyz::ConvexUnionGeometry::box(mthz::Vec3(-100, -1, -100), 200, 1, 200, m), RigidBody::FIXED);
	pin_body(g);
	SplitMix64 rng(seed);
	auto kick = [&](RigidBody* b) {
		b->setVel(mthz::Vec3(rng.uniform(1, 2), 0, rng.uniform(-0.5, 0.5)));
		b->setAngVel(mthz::Vec3(0, rng.uniform(2, 3), 0));
		pin_body(b);
	};
	// attach points are given in the creation frame, i.e. as world points
	RigidBody* a1 = E.createRigidBody(phyz::ConvexUnionGeometry::box(mthz::Vec3(-0.5, 8.5, -0.5), 1, 1, 1, m), RigidBody::FIXED);
	pin_body(a1);
	RigidBody* p1 = E.createRigidBody(phyz::ConvexUnionGeometry::box(mthz::Vec3(-0.2, 7.2, -0.2), 0.4, 1.2, 0.4, m));
	RigidBody* p2 = E.createRigidBody(phyz::ConvexUnionGeometry::box(mthz::Vec3(-0.2, 5.9, -0.2), 0.4, 1.2, 0.4, m));
	kick(p1);
	kick(p2);
	mthz::Vec3 down(0, -1, 0);
	record_joint<phyz::BallSocket>(E, E.addBallSocketConstraint(a1, p1, mthz::Vec3(0, 8.45, 0), mthz::Vec3(0, 8.45, 0)), 3, a1, p1);
	record_joint<phyz::BallSocket>(E, E.addBallSocketConstraint(p1, p2, mthz::Vec3(0, 7.15, 0), mthz::Vec3(0, 7.15, 0)), 3, p1, p2);
	record_joint<phyz::ConeRotationLimit>(E, E.addConeLimitConstraint(a1, p1, down, down, 0.15), 8, a1, p1);
	record_joint<phyz::ConeRotationLimit>(E, E.addConeLimitConstraint(p1, p2, down, down, 0.12), 8, p1, p2);
	RigidBody* a2 = E.createRigidBody(phyz::ConvexUnionGeometry::box(mthz::Vec3(3.5, 4.5, -0.5), 1, 1, 1, m), RigidBody::FIXED);
	pin_body(a2);
	RigidBody* t1 = E.createRigidBody(phyz::ConvexUnionGeometry::box(mthz::Vec3(3.4, 3.2, -0.6), 1.2, 1.2, 1.2, m));
	kick(t1);
	mthz::Vec3 up(0, 1, 0);
	record_joint<phyz::Hinge>(E, E.addHingeConstraint(a2, t1, mthz::Vec3(4, 4.45, 0), mthz::Vec3(4, 4.45, 0), up, up), 5, a2, t1);
	record_joint<phyz::TwistLimit>(E, E.addTwistLimitConstraint(a2, t1, up, -0.3, 0.4), 9, a2, t1);
	Out o(arg_s(argc, argv, "--out", "out"));
	o.meta("seed", (double)seed);
	run_trace(E, cfg, steps, arg_steps(argc, argv, "--record"), o);
	o.finish();
	return 0;
}

// powered: a motor on a hinge (a pendulum arm under a fixed anchor, angle
// limits -1.0 and 1.2) driven through constant torque, target velocity,
// target position and off; a motor on a sliding hinge at a target velocity
// throughout (no limits); a piston on a slider (a block beside a fixed
// anchor, slide limits -0.4 and 0.9) driven through target velocity, target
// position, constant force and off.  The powered-mode changes are applied
// through the engine's setters before their steps and dumped as
// drive_events rows (step, joint, mode, max, target; mode 0 off, 1 constant
// torque / force, 2 target velocity, 3 target position).
static phyz::PersistentConstraint* find_constraint(PhysicsEngine& E, phyz::ConstraintID id) {
	for (auto* x : E.constraint_map[id.getID()]->constraints)
		if (x->getId() == id) return x;
	return nullptr;
}
static int cmd_powered(int argc, char** argv) {
	int steps = (int)arg_d(argc, argv, "--steps", 200);
	SceneCfg cfg;
	cfg.hol_iters = (int)arg_d(argc, argv, "--hol", 1);
	PhysicsEngine E;
	configure(E, cfg);
	phyz::Material m = c0_material();
	RigidBody* g = E.createRigidBody(phyz::ConvexUnionGeometry::box(mthz::Vec3(-100, -1, -100), 200, 1, 200, m), RigidBody::FIXED);
	pin_body(g);
	mthz::Vec3 z(0, 0, 1), x(1, 0, 0);
	// motor on a hinge
	RigidBody* a1 = E.createRigidBody(phyz::ConvexUnionGeometry::box(mthz::Vec3(-0.5, 6.1, -0.5), 1, 1, 1, m), RigidBody::FIXED);
	RigidBody* arm = E.createRigidBody(phyz::ConvexUnionGeometry::box(mthz::Vec3(-0.15, 4.6, -0.15), 0.3, 1.4, 0.3, m));
	pin_body(a1);
	pin_body(arm);
	phyz::ConstraintID h1 = E.addHingeConstraint(a1, arm, mthz::Vec3(0, 6, 0), mthz::Vec3(0, 6, 0), z, z);
	record_joint<phyz::Hinge>(E, h1, 5, a1, arm);
	phyz::ConstraintID m1 = E.addMotorConstraint(h1, -1.0, 1.2);
	auto* mp1 = static_cast<phyz::Motor*>(find_constraint(E, m1));
	g_joints.push_back({10, a1, arm, mp1, &mp1->motor_constraint, static_cast<phyz::Hinge*>(find_constraint(E, h1))->b2_rot_axis_body_space});
	const int jm1 = (int)g_joints.size() - 1;
	// motor on a sliding hinge
	RigidBody* a2 = E.createRigidBody(phyz::ConvexUnionGeometry::box(mthz::Vec3(3.5, 6.1, -0.5), 1, 1, 1, m), RigidBody::FIXED);
	RigidBody* sw = E.createRigidBody(phyz::ConvexUnionGeometry::box(mthz::Vec3(3.85, 4.6, -0.15), 0.3, 1.4, 0.3, m));
	pin_body(a2);
	pin_body(sw);
	phyz::ConstraintID h2 = E.addSlidingHingeConstraint(a2, sw, mthz::Vec3(4, 6, 0), z);
	record_joint<phyz::SlidingHinge>(E, h2, 4, a2, sw);
	phyz::ConstraintID m2 = E.addMotorConstraint(h2, -1e9, 1e9);
	auto* mp2 = static_cast<phyz::Motor*>(find_constraint(E, m2));
	g_joints.push_back({10, a2, sw, mp2, &mp2->motor_constraint, static_cast<phyz::SlidingHinge*>(find_constraint(E, h2))->b2_slide_axis_body_space});
	const int jm2 = (int)g_joints.size() - 1;
	// piston on a slider
	RigidBody* a3 = E.createRigidBody(phyz::ConvexUnionGeometry::box(mthz::Vec3(7.5, 2.5, -0.5), 1, 1, 1, m), RigidBody::FIXED);
	RigidBody* blk = E.createRigidBody(phyz::ConvexUnionGeometry::box(mthz::Vec3(8.7, 2.7, -0.3), 0.6, 0.6, 0.6, m));
	pin_body(a3);
	pin_body(blk);
	phyz::ConstraintID s1 = E.addSliderConstraint(a3, blk, mthz::Vec3(8.5, 3, 0), x);
	record_joint<phyz::Slider>(E, s1, 7, a3, blk);
	phyz::ConstraintID p1 = E.addPistonConstraint(s1, -0.4, 0.9);
	auto* pp = static_cast<phyz::Piston*>(find_constraint(E, p1));
	g_joints.push_back({12, a3, blk, pp, &pp->slide_limit});
	g_joints.push_back({11, a3, blk, pp, &pp->piston_constraint});
	const int jp = (int)g_joints.size() - 1;
	struct Ev { int step, joint, mode; double max, target; };
	std::vector<Ev> ev = {{0, jm1, 1, 4.0, 0},     {40, jm1, 2, 6.0, -2.0}, {80, jm1, 3, 5.0, 0.3},   {130, jm1, 0, 0, 0},
	                      {0, jm2, 2, 3.0, 1.5},   {0, jp, 2, 40.0, 2.0},   {70, jp, 3, 30.0, -0.1},  {120, jp, 1, -20.0, 0},
	                      {160, jp, 0, 0, 0}};
	g_pre_step = [&](PhysicsEngine& e, int k) {
		for (const Ev& v : ev) {
			if (v.step != k) continue;
			phyz::ConstraintID id = v.joint == jm1 ? m1 : (v.joint == jm2 ? m2 : p1);
			bool motor = v.joint != jp;
			switch (v.mode) {
			case 0: motor ? e.setMotorOff(id) : e.setPistonOff(id); break;
			case 1: motor ? e.setMotorConstantTorque(id, v.max) : e.setPistonConstantForce(id, v.max); break;
			case 2: motor ? e.setMotorTargetVelocity(id, v.max, v.target) : e.setPistonTargetVelocity(id, v.max, v.target); break;
			default: motor ? e.setMotorTargetPosition(id, v.max, v.target) : e.setPistonTargetPosition(id, v.max, v.target); break;
			}
		}
	};
	Out o(arg_s(argc, argv, "--out", "out"));
	std::vector<double> evd;
	for (const Ev& v : ev) evd.insert(evd.end(), {(double)v.step, (double)v.joint, (double)v.mode, v.max, v.target});
	o.f64("drive_events", evd, {ev.size(), 5});
	run_trace(E, cfg, steps, arg_steps(argc, argv, "--record"), o);
	o.finish();
	return 0;
}

// springs: a box hanging from a fixed anchor by a spring (resting length the
// current distance), a second box hanging from it by a spring of resting
// length 1.0 that lands on the ground, and two boxes on the ground pushed
// apart by a compressed spring; a three-link ball-socket chain from a fixed
// anchor (one holonomic system) whose middle joint has a custom CFM, and a
// hinged block with a custom CFM.
static int cmd_springs(int argc, char** argv) {
	int steps = (int)arg_d(argc, argv, "--steps", 200);
	SceneCfg cfg;
	cfg.hol_iters = (int)arg_d(argc, argv, "--hol", 1);
	PhysicsEngine E;
	configure(E, cfg);
	phyz::Material m = c0_material();
	RigidBody* g = E.createRigidBody(phyz::ConvexUnionGeometry::box(mthz::Vec3(-100, -1, -100), 200, 1, 200, m), RigidBody::FIXED);
	pin_body(g);
	auto spring = [&](RigidBody* b1, RigidBody* b2, mthz::Vec3 p1, mthz::Vec3 p2, double damping, double stiffness, double rest) {
		phyz::ConstraintID id = E.addSpring(b1, b2, p1, p2, damping, stiffness, rest);
		g_joints.push_back({13, b1, b2, find_constraint(E, id), nullptr, mthz::Vec3(rest, 0, 0)});
	};
	RigidBody* a = E.createRigidBody(phyz::ConvexUnionGeometry::box(mthz::Vec3(-0.5, 6.0, -0.5), 1, 1, 1, m), RigidBody::FIXED);
	RigidBody* s1 = E.createRigidBody(phyz::ConvexUnionGeometry::box(mthz::Vec3(-0.3, 3.9, -0.3), 0.6, 0.6, 0.6, m));
	RigidBody* s2 = E.createRigidBody(phyz::ConvexUnionGeometry::box(mthz::Vec3(-0.2, 2.2, 0.0), 0.4, 0.4, 0.4, m));
	pin_body(a);
	pin_body(s1);
	pin_body(s2);
	s1->setVel(mthz::Vec3(0.4, 0, 0.2));
	spring(a, s1, mthz::Vec3(0, 6.0, 0), mthz::Vec3(0.1, 4.5, 0), 2.0, 60.0, -1);
	spring(s1, s2, mthz::Vec3(0, 3.9, 0), mthz::Vec3(0, 2.6, 0.2), 0.5, 40.0, 1.0);
	RigidBody* f1 = E.createRigidBody(phyz::ConvexUnionGeometry::box(mthz::Vec3(3.0, 0.0, -0.4), 0.8, 0.8, 0.8, m));
	RigidBody* f2 = E.createRigidBody(phyz::ConvexUnionGeometry::box(mthz::Vec3(4.2, 0.0, -0.4), 0.8, 0.8, 0.8, m));
	pin_body(f1);
	pin_body(f2);
	spring(f1, f2, mthz::Vec3(3.8, 0.4, 0), mthz::Vec3(4.2, 0.4, 0), 0.2, 80.0, 1.2);
	// custom CFM on joints
	RigidBody* c0 = E.createRigidBody(phyz::ConvexUnionGeometry::box(mthz::Vec3(-6.5, 7.5, -0.5), 1, 1, 1, m), RigidBody::FIXED);
	pin_body(c0);
	RigidBody* prev = c0;
	double y = 7.5;
	for (int i = 0; i < 3; i++) {
		RigidBody* l = E.createRigidBody(phyz::ConvexUnionGeometry::box(mthz::Vec3(-6.2 + 0.5 * i, y - 1.0, -0.2), 0.4, 1.0, 0.4, m));
		pin_body(l);
		l->setVel(mthz::Vec3(1.0, 0, 0.5));
		phyz::ConstraintID id = E.addBallSocketConstraint(prev, l, mthz::Vec3(-6.0 + 0.5 * i, y, 0), mthz::Vec3(-6.0 + 0.5 * i, y, 0));
		record_joint<phyz::BallSocket>(E, id, 3, prev, l);
		if (i == 1) {
			E.setConstraintUseCustomCFM(id, 1e-3);
			g_joints.back().ccfm = 1e-3;
		}
		prev = l;
		y -= 1.0;
	}
	RigidBody* ha = E.createRigidBody(phyz::ConvexUnionGeometry::box(mthz::Vec3(7.5, 4.5, -0.5), 1, 1, 1, m), RigidBody::FIXED);
	RigidBody* hb = E.createRigidBody(phyz::ConvexUnionGeometry::box(mthz::Vec3(7.4, 3.2, -0.6), 1.2, 1.2, 1.2, m));
	pin_body(ha);
	pin_body(hb);
	hb->setAngVel(mthz::Vec3(0, 2.0, 0));
	hb->setVel(mthz::Vec3(0, 0, 1.0));
	phyz::ConstraintID hid = E.addHingeConstraint(ha, hb, mthz::Vec3(8, 4.45, 0), mthz::Vec3(8, 4.45, 0), mthz::Vec3(0, 1, 0), mthz::Vec3(0, 1, 0));
	record_joint<phyz::Hinge>(E, hid, 5, ha, hb);
	E.setConstraintUseCustomCFM(hid, 5e-4);
	g_joints.back().ccfm = 5e-4;
	Out o(arg_s(argc, argv, "--out", "out"));
	run_trace(E, cfg, steps, arg_steps(argc, argv, "--record"), o);
	o.finish();
	return 0;
}

// sleep: sleeping enabled (the engine's default parameters, every body's
// sleep_disabled pinned to false): a 2x2x2 cube pile and a lone cube settle
// and fall asleep, a cube launched along the ground slides to rest and falls
// asleep later, and a sphere dropped from high above lands on a sleeping
// column of three cubes and wakes it; a waking action on one cube of the
// sleeping pile (before step --wake-step, default 250) wakes the pile.
// Every step's sleep state is recorded (asleep, sleep_ready_counter,
// non_sleepy_tick_count, history length).
static int cmd_sleep(int argc, char** argv) {
	int steps = (int)arg_d(argc, argv, "--steps", 420);
	uint64_t seed = (uint64_t)arg_d(argc, argv, "--seed", 5150);
	double drop = arg_d(argc, argv, "--drop", 24.0);
	SceneCfg cfg;
	PhysicsEngine E;
	configure(E, cfg);
	E.setSleepingEnabled(true);
	phyz::Material m = c0_material();
	auto pin = [](RigidBody* r) {
		r->setNoCollision(false);
		r->setSleepDisabled(false);
	};
	RigidBody* g = E.createRigidBody(phyz::ConvexUnionGeometry::box(mthz::Vec3(-100, -1, -100), 200, 1, 200, m), RigidBody::FIXED);
	pin(g);
	SplitMix64 rng(seed);
	for (int j = 0; j < 2; j++)
		for (int i = 0; i < 2; i++)
			for (int k = 0; k < 2; k++) {
				double cx = 1.1 * i + rng.uniform(-0.05, 0.05), cy = 0.6 + 1.1 * j + rng.uniform(-0.05, 0.05);
				double cz = 1.1 * k + rng.uniform(-0.05, 0.05), yaw = rng.uniform(0.0, 2.0 * 3.14159265358979323846);
				RigidBody* r = E.createRigidBody(phyz::ConvexUnionGeometry::box(mthz::Vec3(cx - 0.5, cy - 0.5, cz - 0.5), 1, 1, 1, m));
				r->setOrientation(mthz::Quaternion(yaw, mthz::Vec3(0, 1, 0)));
				pin(r);
			}
	RigidBody* lone = E.createRigidBody(phyz::ConvexUnionGeometry::box(mthz::Vec3(-5.5, 0.05, -0.5), 1, 1, 1, m));
	pin(lone);
	RigidBody* slider = E.createRigidBody(phyz::ConvexUnionGeometry::box(mthz::Vec3(-0.5, 0.0, -6.5), 1, 1, 1, m));
	slider->setVel(mthz::Vec3(2.5, 0, 0));
	pin(slider);
	for (int j = 0; j < 3; j++) {
		RigidBody* r = E.createRigidBody(phyz::ConvexUnionGeometry::box(mthz::Vec3(5.5, 0.02 + 1.02 * j, -0.5), 1, 1, 1, m));
		pin(r);
	}
	RigidBody* ball = E.createRigidBody(phyz::ConvexUnionGeometry::sphere(mthz::Vec3(6.0 + 0.1, drop, 0.05), 0.5, m));
	pin(ball);
	Out o(arg_s(argc, argv, "--out", "out"));
	o.meta("seed", (double)seed);
	g_sleep = true;
	g_wake_step = (int)arg_d(argc, argv, "--wake-step", 250);
	g_wake_body = (int)arg_d(argc, argv, "--wake-body", 3);
	o.meta("wake_step", g_wake_step);
	o.meta("wake_body", g_wake_body);
	run_trace(E, cfg, steps, arg_steps(argc, argv, "--record"), o);
	o.finish();
	return 0;
}

// hulls: irregular convex hulls (points and faces read from a text file: the
// count, then per hull "n_pts n_faces", the points, and each face as its
// vertex count and indices) built as createRigidBody(ConvexUnionGeometry(
// Polyhedron(points, faces))) does (RigidBody.cpp:13-38), mixed with spheres
// and cylinders created centred on the origin, in an nx x ny x nz grid of
// pitch `spacing` above the ground box; kinds cycle hull, hull, sphere,
// cylinder; each body gets a random orientation (angle U[0, 2 pi) about a
// normalised U[-1, 1]^3 axis) and velocity U[-0.5, 0.5]^3.  Spheres r U[0.3,
// 0.7], cylinders r U[0.2, 0.5], h U[0.5, 1.5]; default material.
struct HullIn {
	std::vector<mthz::Vec3> pts;
	std::vector<std::vector<uint32_t>> faces;
};
static std::vector<HullIn> read_hulls(const std::string& path) {
	FILE* f = fopen(path.c_str(), "r");
	if (!f) { perror(path.c_str()); exit(2); }
	int k = 0;
	if (fscanf(f, "%d", &k) != 1) exit(2);
	std::vector<HullIn> out(k);
	for (auto& h : out) {
		int np = 0, nf = 0;
		if (fscanf(f, "%d %d", &np, &nf) != 2) exit(2);
		for (int i = 0; i < np; i++) {
			char b[3][64];
			if (fscanf(f, "%63s %63s %63s", b[0], b[1], b[2]) != 3) exit(2);
			h.pts.push_back(mthz::Vec3(strtod(b[0], nullptr), strtod(b[1], nullptr), strtod(b[2], nullptr)));
		}
		for (int i = 0; i < nf; i++) {
			int m = 0;
			if (fscanf(f, "%d", &m) != 1) exit(2);
			std::vector<uint32_t> fv(m);
			for (auto& x : fv)
				if (fscanf(f, "%u", &x) != 1) exit(2);
			h.faces.push_back(fv);
		}
	}
	fclose(f);
	return out;
}

static int cmd_hulls(int argc, char** argv) {
	int nx = (int)arg_d(argc, argv, "--n", 3);
	int ny = (int)arg_d(argc, argv, "--ny", nx);
	int steps = (int)arg_d(argc, argv, "--steps", 10);
	double spacing = arg_d(argc, argv, "--spacing", 1.8);
	uint64_t seed = (uint64_t)arg_d(argc, argv, "--seed", 31337);
	std::vector<HullIn> in = read_hulls(arg_s(argc, argv, "--in", "hull_input.txt"));
	SceneCfg cfg;
	PhysicsEngine E;
	configure(E, cfg);
	phyz::Material m = phyz::Material::default_material();
	RigidBody* g = E.createRigidBody(phyz::ConvexUnionGeometry::box(mthz::Vec3(-100, -1, -100), 200, 1, 200, m), RigidBody::FIXED);
	pin_body(g);
	SplitMix64 rng(seed);
	std::vector<int64_t> hull_of{-1};  // per body: the input hull it was built from
	int idx = 0, next_hull = 0;
	const double PI_ = 3.14159265358979323846;
	for (int j = 0; j < ny; j++)
		for (int i = 0; i < nx; i++)
			for (int k = 0; k < nx; k++) {
				mthz::Vec3 c(spacing * (i - 0.5 * (nx - 1)) + rng.uniform(-0.05, 0.05), 1.0 + spacing * j + rng.uniform(-0.05, 0.05),
				             spacing * (k - 0.5 * (nx - 1)) + rng.uniform(-0.05, 0.05));
				double th = rng.uniform(0.0, 2.0 * PI_);
				mthz::Vec3 ax(rng.uniform(-1.0, 1.0), rng.uniform(-1.0, 1.0), rng.uniform(-1.0, 1.0));
				mthz::Vec3 v(rng.uniform(-0.5, 0.5), rng.uniform(-0.5, 0.5), rng.uniform(-0.5, 0.5));
				double r = rng.uniform(0.3, 0.7), cr = rng.uniform(0.2, 0.5), h = rng.uniform(0.5, 1.5);
				int kind = idx++ % 4;
				RigidBody* b;
				if (kind < 2) {
					const HullIn& hi = in[next_hull % in.size()];
					hull_of.push_back(next_hull % (int)in.size());
					next_hull++;
					b = E.createRigidBody(phyz::ConvexUnionGeometry(phyz::ConvexPrimitive((const phyz::ConvexGeometry&)phyz::Polyhedron(hi.pts, hi.faces), m)));
				} else if (kind == 2) {
					hull_of.push_back(-1);
					b = E.createRigidBody(phyz::ConvexUnionGeometry::sphere(mthz::Vec3(0, 0, 0), r, m));
				} else {
					hull_of.push_back(-1);
					b = E.createRigidBody(phyz::ConvexUnionGeometry::cylinder(mthz::Vec3(0, -h / 2.0, 0), cr, h, m));
				}
				b->setToPosition(c);
				b->setOrientation(mthz::Quaternion(th, ax.normalize()));
				b->setVel(v);
				pin_body(b);
			}
	Out o(arg_s(argc, argv, "--out", "out"));
	o.meta("seed", (double)seed);
	std::vector<double> ip;
	std::vector<int64_t> ipo{0}, ifo{0}, ifv, ifoff{0};
	for (auto& h : in) {
		for (auto& p : h.pts) ip.insert(ip.end(), {p.x, p.y, p.z});
		ipo.push_back((int64_t)ip.size() / 3);
		for (auto& fv : h.faces) {
			for (uint32_t x : fv) ifv.push_back(x);
			ifoff.push_back((int64_t)ifv.size());
		}
		ifo.push_back((int64_t)ifoff.size() - 1);
	}
	o.f64("input_hull_points", ip, {ip.size() / 3, 3});
	o.i64("input_hull_point_off", ipo, {ipo.size()});
	o.i64("input_hull_face_off", ifo, {ifo.size()});
	o.i64("input_hull_face_idx", ifoff, {ifoff.size()});
	o.i64("input_hull_face_verts", ifv, {ifv.size()});
	o.i64("input_hull_of_body", hull_of, {hull_of.size()});
	run_trace(E, cfg, steps, arg_steps(argc, argv, "--record"), o);
	dump_rays(E, o, (int)arg_d(argc, argv, "--rays", 0), seed);
	o.finish();
	return 0;
}

// time: single-threaded wall time of the reference's own timeStep on a C0 pile.
static int cmd_time(int argc, char** argv) {
	int n = (int)arg_d(argc, argv, "--n", 10);
	int steps = (int)arg_d(argc, argv, "--steps", 300);
	int warm = (int)arg_d(argc, argv, "--warmup", 10);
	double budget = arg_d(argc, argv, "--budget-s", 1e30);
	uint64_t seed = (uint64_t)arg_d(argc, argv, "--seed", 12345);
	SceneCfg cfg;
	PhysicsEngine E;
	configure(E, cfg);
	build_pile(E, n, n, n, seed, 1.1, 0.05);
	int dyn = 0;
	for (RigidBody* b : E.bodies) dyn += b->getMovementType() == RigidBody::DYNAMIC;
	for (int k = 0; k < warm; k++) E.timeStep();
	std::vector<double> t;
	double total = 0;
	for (int k = 0; k < steps; k++) {
		auto a = std::chrono::steady_clock::now();
		E.timeStep();
		auto b = std::chrono::steady_clock::now();
		double s = std::chrono::duration<double>(b - a).count();
		t.push_back(s);
		total += s;
		if (total > budget) break;
	}
	std::vector<double> st = t;
	std::sort(st.begin(), st.end());
	double p50 = st[st.size() / 2], p99 = st[std::min(st.size() - 1, (size_t)(0.99 * st.size()))];
	printf("{\"bodies\": %d, \"steps\": %zu, \"warmup\": %d, \"wall_s\": %.6f, \"body_steps_per_s\": %.3f, \"p50_ms\": %.3f, \"p99_ms\": %.3f}\n",
	       dyn, t.size(), warm, total, dyn * (double)t.size() / total, p50 * 1e3, p99 * 1e3);
	return 0;
}

int main(int argc, char** argv) {
	if (argc < 2) {
		fprintf(stderr, "usage: cp_ref_harness pile|time [options]\n");
		return 2;
	}
	std::string cmd = argv[1];
	if (cmd == "pile") return cmd_pile(argc, argv);
	if (cmd == "chain") return cmd_chain(argc, argv);
	if (cmd == "terrain") return cmd_terrain(argc, argv);
	if (cmd == "time") return cmd_time(argc, argv);
	if (cmd == "hulls") return cmd_hulls(argc, argv);
	if (cmd == "sleep") return cmd_sleep(argc, argv);
	if (cmd == "joints") return cmd_joints(argc, argv);
	if (cmd == "limits") return cmd_limits(argc, argv);
	if (cmd == "powered") return cmd_powered(argc, argv);
	if (cmd == "springs") return cmd_springs(argc, argv);
	fprintf(stderr, "unknown command %s\n", cmd.c_str());
	return 2;
}
