// cp_oracle — FP64 CPU restatement of ConstraintPhysics' per-step contact
// pipeline.  TEST INFRASTRUCTURE ONLY (see cp_oracle.h).  Scalar code that
// keeps the reference's data flow and floating-point operation order; compiled
// with -ffp-contract=off.  Reference citations are file:line in
// NathanMacLeod/ConstraintPhysics.
//
// Parity status (pinned by tests/test_oracle.py against reference fixtures):
//   world geometry / AABBs, broadphase set, SAT poly-poly + clipping + cull,
//   contact persistence, island order, row build, PGS, integration: pinned
//   bit-for-bit on C0-pattern box piles.
#include "cp_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <deque>
#include <limits>
#include <map>
#include <set>
#include <climits>
#include <memory>
#include <unordered_map>
#include <vector>

namespace {

// ---------------------------------------------------------------- math (Math/src)
struct V3 {
	double x = 0, y = 0, z = 0;
	V3() {}
	V3(double a, double b, double c) : x(a), y(b), z(c) {}
	V3 operator+(V3 v) const { return V3(x + v.x, y + v.y, z + v.z); }
	V3 operator-(V3 v) const { return V3(x - v.x, y - v.y, z - v.z); }
	V3 operator-() const { return V3(-x, -y, -z); }
	V3 operator*(double d) const { return V3(x * d, y * d, z * d); }  // Vec3.cpp:4-9
	V3 operator/(double d) const { return V3(x / d, y / d, z / d); }
	void operator+=(V3 v) { x += v.x; y += v.y; z += v.z; }
	void operator-=(V3 v) { x -= v.x; y -= v.y; z -= v.z; }
	double dot(V3 v) const { return x * v.x + y * v.y + z * v.z; }
	V3 cross(V3 v) const { return V3(y * v.z - z * v.y, z * v.x - x * v.z, x * v.y - y * v.x); }
	double magSqrd() const { return x * x + y * y + z * z; }
	double mag() const { return std::sqrt(x * x + y * y + z * z); }
	V3 normalize() const {  // Vec3.h:38-41
		double m = mag();
		return (m == 0) ? V3(0, 0, 0) : V3(x / m, y / m, z / m);
	}
	bool operator!=(V3 v) const { return x != v.x || y != v.y || z != v.z; }
	// Vec3.h:53-57
	void perpBasis(V3* u, V3* w) const {
		V3 a1(1, 0, 0), a2(0, 1, 0);
		*u = (std::abs(dot(a1)) < std::abs(dot(a2))) ? cross(a1).normalize() : cross(a2).normalize();
		*w = cross(*u).normalize();
	}
};
static inline V3 operator*(double d, V3 v) { return V3(v.x * d, v.y * d, v.z * d); }

struct M3 {
	double v[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
	V3 operator*(V3 w) const {  // Mat3.h:70-74
		return V3(v[0][0] * w.x + v[0][1] * w.y + v[0][2] * w.z, v[1][0] * w.x + v[1][1] * w.y + v[1][2] * w.z,
		          v[2][0] * w.x + v[2][1] * w.y + v[2][2] * w.z);
	}
	M3 operator*(const M3& m) const {  // Mat3.h:95-106
		M3 o;
		for (int i = 0; i < 3; i++)
			for (int j = 0; j < 3; j++)
				for (int k = 0; k < 3; k++) o.v[i][j] += v[i][k] * m.v[k][j];
		return o;
	}
	M3 operator+(const M3& m) const {
		M3 o;
		for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) o.v[i][j] = v[i][j] + m.v[i][j];
		return o;
	}
	M3 operator-(const M3& m) const {
		M3 o;
		for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) o.v[i][j] = v[i][j] - m.v[i][j];
		return o;
	}
	double det() const {  // Mat3.h:52-55
		double a = v[0][0], b = v[0][1], c = v[0][2], d = v[1][0], e = v[1][1], f = v[1][2], g = v[2][0], h = v[2][1], i = v[2][2];
		return a * (e * i - f * h) - b * (d * i - f * g) + c * (d * h - e * g);
	}
	M3 inverse() const {  // Mat3.h:57-68
		M3 o;
		double a = v[0][0], b = v[0][1], c = v[0][2], d = v[1][0], e = v[1][1], f = v[1][2], g = v[2][0], h = v[2][1], i = v[2][2];
		double dt = a * (e * i - f * h) - b * (d * i - f * g) + c * (d * h - e * g);
		o.v[0][0] = (e * i - f * h); o.v[0][1] = (c * h - b * i); o.v[0][2] = (b * f - c * e);
		o.v[1][0] = (f * g - d * i); o.v[1][1] = (a * i - c * g); o.v[1][2] = (c * d - a * f);
		o.v[2][0] = (d * h - e * g); o.v[2][1] = (b * g - a * h); o.v[2][2] = (a * e - b * d);
		for (int r = 0; r < 3; r++) for (int s = 0; s < 3; s++) o.v[r][s] /= dt;
		return o;
	}
	static M3 cross_mat(V3 w) {  // Mat3.h:43-49
		M3 o;
		o.v[0][0] = 0; o.v[0][1] = -w.z; o.v[0][2] = w.y;
		o.v[1][0] = w.z; o.v[1][1] = 0; o.v[1][2] = -w.x;
		o.v[2][0] = -w.y; o.v[2][1] = w.x; o.v[2][2] = 0;
		return o;
	}
};
static inline M3 operator*(double d, const M3& m) {  // Mat3.cpp:4-12
	M3 o;
	for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) o.v[i][j] = d * m.v[i][j];
	return o;
}

struct Q {
	double r = 1, i = 0, j = 0, k = 0;
	Q() {}
	Q(double a, double b, double c, double d) : r(a), i(b), j(c), k(d) {}
	Q(double theta, V3 axis) {  // Quaternion.h:26-33
		axis = axis.normalize();
		r = std::cos(theta / 2.0);
		double s = std::sin(theta / 2.0);
		i = axis.x * s; j = axis.y * s; k = axis.z * s;
	}
	V3 rotate(V3 v) const {  // Quaternion.h:51-55
		return V3(v.x * (r * r + i * i - j * j - k * k) + v.y * (2 * i * j - 2 * r * k) + v.z * (2 * i * k + 2 * r * j),
		          v.x * (2 * i * j + 2 * r * k) + v.y * (r * r - i * i + j * j - k * k) + v.z * (2 * j * k - 2 * r * i),
		          v.x * (2 * i * k - 2 * r * j) + v.y * (2 * j * k + 2 * r * i) + v.z * (r * r - i * i - j * j + k * k));
	}
	M3 rot() const {  // Quaternion.h:61-67
		M3 m;
		m.v[0][0] = r * r + i * i - j * j - k * k; m.v[0][1] = 2 * i * j - 2 * r * k; m.v[0][2] = 2 * i * k + 2 * r * j;
		m.v[1][0] = 2 * i * j + 2 * r * k; m.v[1][1] = r * r - i * i + j * j - k * k; m.v[1][2] = 2 * j * k - 2 * r * i;
		m.v[2][0] = 2 * i * k - 2 * r * j; m.v[2][1] = 2 * j * k + 2 * r * i; m.v[2][2] = r * r - i * i - j * j + k * k;
		return m;
	}
	Q conj() const { return Q(r, -i, -j, -k); }
	Q operator*(Q q) const {  // Quaternion.h:73-78
		return Q(r * q.r - i * q.i - j * q.j - k * q.k, r * q.i + i * q.r + j * q.k - k * q.j, r * q.j - i * q.k + j * q.r + k * q.i,
		         r * q.k + i * q.j - j * q.i + k * q.r);
	}
	Q normalize() const {  // Quaternion.h:88-91
		double m = std::sqrt(r * r + i * i + j * j + k * k);
		return Q(r / m, i / m, j / m, k / m);
	}
	bool operator!=(Q q) const { return r != q.r || i != q.i || j != q.j || k != q.k; }
};

// Fixed-size row algebra with NMat's accumulation order (NMat.h:50-60, :91-103).
template <int R, int C>
struct NM {
	double v[R][C] = {};
};
template <int R, int C, int X>
static NM<R, X> mul(const NM<R, C>& a, const NM<C, X>& b) {
	NM<R, X> o;
	for (int i = 0; i < R; i++)
		for (int j = 0; j < X; j++) {
			o.v[i][j] = 0;
			for (int k = 0; k < C; k++) o.v[i][j] += a.v[i][k] * b.v[k][j];
		}
	return o;
}
template <int R, int C>
static void mulv(const NM<R, C>& a, const double* x, double* out) {
	for (int i = 0; i < R; i++) {
		out[i] = 0;
		for (int j = 0; j < C; j++) out[i] += x[j] * a.v[i][j];
	}
}
template <int R, int C>
static NM<C, R> tr(const NM<R, C>& a) {
	NM<C, R> o;
	for (int i = 0; i < R; i++) for (int j = 0; j < C; j++) o.v[j][i] = a.v[i][j];
	return o;
}
template <int R, int C>
static NM<R, C> neg(const NM<R, C>& a) {
	NM<R, C> o;
	for (int i = 0; i < R; i++) for (int j = 0; j < C; j++) o.v[i][j] = -a.v[i][j];
	return o;
}
static NM<3, 3> crossNM(V3 v) {  // NMat.cpp:62-70 (untouched entries stay 0)
	NM<3, 3> o;
	o.v[0][1] = -v.z; o.v[0][2] = v.y;
	o.v[1][0] = v.z; o.v[1][2] = -v.x;
	o.v[2][0] = -v.y; o.v[2][1] = v.x;
	return o;
}
// Constraint::aInvMass (ConstraintSolver.cpp:215-225)
static NM<6, 6> invMass6(double im, const M3& it) {
	NM<6, 6> o;
	for (int i = 0; i < 3; i++) o.v[i][i] = 1.0 * im;
	for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) o.v[3 + i][3 + j] = it.v[i][j];
	// idenMat<3>() * inv_mass multiplies the zeros too; keep them exact zeros.
	for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) if (i != j) o.v[i][j] = 0.0 * im;
	return o;
}
static bool isZero(const double* v, int n) {  // NVec.h:12-19
	for (int i = 0; i < n; i++) if (std::abs(v[i]) > 0.00000001) return false;
	return true;
}

// Reference constants (CollisionDetect.h:20-26).
static const double REF_PI = 3.14159265358979323846;
static const double TOL_ANG = 1 * REF_PI / 180;
static const double COS_TOL = 1 - std::cos(TOL_ANG);
static const double SIN_TOL = std::sin(TOL_ANG);

// ---------------------------------------------------------------- geometry
struct Hull {
	std::vector<V3> ref_pts;
	std::vector<std::vector<int>> faces;
	std::vector<int> surf_id, ncalc, ndir;
	std::vector<V3> gnorm;  // body frame
	std::vector<int> gred, gmin_pid, gmax_pid, gref_idx;
	std::vector<double> gmin, gmax, gref_val;
	std::vector<std::pair<int, int>> arcs, edges;
	std::vector<std::vector<int>> adj_faces, adj_edges;
	// world copy
	std::vector<V3> pts;
	std::vector<V3> wnorm;
};

struct Mesh;

struct Body {
	int id;
	int movement;  // 0 dynamic, 1 fixed
	double mass, inv_mass;
	M3 I_ref, I_ref_inv, I, I_inv;
	double restitution, mu_k, mu_s;
	V3 com, vel, ang_vel, pvel, pang;
	Q q;
	Hull hull;
	int type = 0;        // 0 polyhedron, 1 sphere, 3 cylinder (ConvexGeometryType, ConvexPrimitive.h:36), 5 static mesh
	std::shared_ptr<Mesh> mesh;  // static triangle mesh (RigidBody::mesh / reference_mesh, RigidBody.h)
	V3 sph_ref;          // sphere centre in the body frame (Sphere::recomputeFromReference)
	double radius = 0;
	V3 sph_c;            // world centre
	// cylinder (ConvexPrimitive.h:159-210): body-frame reference and world copy;
	// 14 Gauss vertices (caps, then the 12-gon ring), 8-gon cap polygons
	V3 cyl_c_ref, cyl_ax_ref, cyl_gv_ref[14], cyl_top_ref[8], cyl_bot_ref[8];
	double cyl_r = 0, cyl_h = 0;
	V3 cyl_c, cyl_ax, cyl_gv[14], cyl_top[8], cyl_bot[8];
	double aabb[6];
	V3 prev_com;
	Q prev_q;
	// sleeping (RigidBody.h:88-124): asleep, sleep_ready_counter,
	// non_sleepy_tick_count, history of (vel, ang_vel), recievedWakingAction,
	// sleep_disabled
	bool asleep = false, wake_req = false, sleep_disabled = false;
	double sleep_counter = 0;
	int sleep_ticks = 0;
	std::deque<std::pair<V3, V3>> history;
	bool dynamic() const { return movement == 0; }
	bool getAsleep() const { return asleep && movement == 0; }  // RigidBody.cpp:181-183
	double invMass() const { return movement == 0 ? 1.0 / mass : 0; }       // RigidBody.cpp:162-164
	M3 invTensor() const { return movement == 0 ? I_inv : M3(); }          // RigidBody.cpp:170-172
	V3 getVel() const { return movement == 1 ? V3() : vel; }
	V3 getAngVel() const { return movement == 1 ? V3() : ang_vel; }
};

// Surface::normal (ConvexPrimitive.cpp:815-821), evaluated on world points.
static V3 faceNormal(const Hull& h, int f) {
	const std::vector<int>& pi = h.faces[f];
	V3 p1 = h.pts[pi[0]], p2 = h.pts[pi[1]], p3 = h.pts[pi[h.ncalc[f]]];
	return (p2 - p1).cross(p3 - p2).normalize() * (double)h.ndir[f];
}

// ---------------------------------------------------------------- static triangle mesh
// StaticMeshFace (Geometry.h:61-86): world triangle with its edges, Gauss
// region (directions a contact normal may take against this triangle) and
// material.
struct MeshTri {
	V3 n;
	V3 v[3];
	int vid[3];
	V3 ep1[3], ep2[3], eout[3];
	int eid[3];
	std::vector<V3> gauss;
	int concave = 0;
	double rest = 0, mu_k = 0, mu_s = 0;
	double aabb[6];
	int id;
};
// AABBTree<unsigned int> node (AABB_Tree.h:272-350); leaves hold a triangle.
struct BVNode {
	double box[6], tbox[6];
	double cost = 0;
	int parent = -1, left = -1, right = -1, leaf = -1;
};
struct Mesh {
	std::vector<MeshTri> ref, tris;  // reference (body frame) and world triangles
	std::vector<BVNode> nodes;
	int root = -1;
};

static void boxCombine(const double* a, const double* b, double* o) {  // AABB::combine (AABB.h:16-26)
	for (int i = 0; i < 3; i++) {
		o[i] = std::min<double>(a[i], b[i]);
		o[3 + i] = std::max<double>(a[3 + i], b[3 + i]);
	}
}
static double boxSurface(const double* a) {  // AABB::surfaceArea (AABB.h:112-118), as written
	double dx = a[3] - a[0], dy = a[4] - a[1], dz = a[5] - a[2];
	return 2 * (dx * dy) + (dy * dz) + (dz * dx);
}
static double boxLongest(const double* a) {  // AABB::longestDimension (AABB.h:120-122)
	return std::max<double>(a[3] - a[0], std::max<double>(a[4] - a[1], a[5] - a[2]));
}
static void triBox(MeshTri& t) {  // AABB::encapsulatePointCloud (AABB.h:52-70)
	const double inf = std::numeric_limits<double>::infinity();
	double mn[3] = {inf, inf, inf}, mx[3] = {-inf, -inf, -inf};
	for (int i = 0; i < 3; i++) {
		mn[0] = std::min<double>(mn[0], t.v[i].x); mn[1] = std::min<double>(mn[1], t.v[i].y); mn[2] = std::min<double>(mn[2], t.v[i].z);
		mx[0] = std::max<double>(mx[0], t.v[i].x); mx[1] = std::max<double>(mx[1], t.v[i].y); mx[2] = std::max<double>(mx[2], t.v[i].z);
	}
	for (int i = 0; i < 3; i++) { t.aabb[i] = mn[i]; t.aabb[3 + i] = mx[i]; }
}

// StaticMeshGeometry(const MeshInput&) (Geometry.cpp:678-820): triangle
// normals, neighbours through shared edges, ids (triangles 0..T-1, vertices
// T + point index, edges from T + P upwards, three fresh ids per triangle)
// and the Gauss region of each triangle from the convexity of its edges.
static void meshBuild(Mesh& M, int n_pts, const double* pts, int n_tris, const int64_t* idx, const double* mat3) {
	std::vector<V3> P(n_pts);
	for (int i = 0; i < n_pts; i++) P[i] = V3(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
	struct E { int a, b, opp; uint64_t key; };
	struct Node { int p[3]; int nb[3]; E e[3]; V3 n; };
	std::vector<Node> g(n_tris);
	for (int t = 0; t < n_tris; t++) {
		int i1 = (int)idx[3 * t], i2 = (int)idx[3 * t + 1], i3 = (int)idx[3 * t + 2];
		Node& nd = g[t];
		nd.p[0] = i1; nd.p[1] = i2; nd.p[2] = i3;
		nd.n = (P[i2] - P[i1]).cross(P[i3] - P[i1]).normalize();
		int ev[3][3] = {{i1, i2, i3}, {i2, i3, i1}, {i3, i1, i2}};
		for (int k = 0; k < 3; k++) {
			uint64_t mx = (uint64_t)std::max(ev[k][0], ev[k][1]), mn = (uint64_t)std::min(ev[k][0], ev[k][1]);
			nd.e[k] = {ev[k][0], ev[k][1], ev[k][2], ((mx << 32) & 0xFFFFFFFF00000000ull) + mn};
			nd.nb[k] = -1;
		}
	}
	std::unordered_map<uint64_t, int> open;
	for (int t = 0; t < n_tris; t++)
		for (int k = 0; k < 3; k++) {
			auto it = open.find(g[t].e[k].key);
			if (it != open.end()) {
				g[t].nb[k] = it->second;
				for (int q = 0; q < 3; q++)
					if (g[it->second].e[q].key == g[t].e[k].key) g[it->second].nb[q] = t;
				open.erase(it);
			} else {
				open[g[t].e[k].key] = t;
			}
		}
	int next_edge = n_tris + n_pts;
	M.ref.resize(n_tris);
	for (int t = 0; t < n_tris; t++) {
		const Node& nd = g[t];
		MeshTri& tr = M.ref[t];
		tr.n = nd.n;
		for (int k = 0; k < 3; k++) {
			tr.v[k] = P[nd.p[k]];
			tr.vid[k] = nd.p[k] + n_tris;
		}
		tr.rest = mat3[3 * t];
		tr.mu_k = mat3[3 * t + 1];
		tr.mu_s = mat3[3 * t + 2];
		tr.id = t;
		tr.concave = 0;
		for (int k = 0; k < 3; k++) {
			tr.ep1[k] = tr.v[k];
			tr.ep2[k] = tr.v[(k + 1) % 3];
			tr.eout[k] = (tr.ep2[k] - tr.ep1[k]).cross(tr.n).normalize();
			tr.eid[k] = next_edge++;
			if (nd.nb[k] == -1) {
				tr.gauss.push_back(tr.eout[k]);
				continue;
			}
			const Node& o = g[nd.nb[k]];
			int opp = -1;
			for (int q = 0; q < 3; q++)
				if (nd.e[k].a == o.e[q].b && nd.e[k].b == o.e[q].a) opp = o.e[q].opp;
			V3 tip = P[nd.e[k].opp], otip = P[opp];
			const double EPS = 0.00001;
			if ((otip - tip).normalize().dot(nd.n) >= -EPS) {
				tr.concave++;
				if (tr.concave <= 1) tr.gauss.push_back(tr.n);
			} else {
				tr.gauss.push_back(o.n);
			}
		}
		triBox(tr);
	}
}

// StaticMeshFace::getTransformed (Geometry.cpp:827-847), centre of rotation 0.
static MeshTri triTransform(const MeshTri& r, const M3& R, V3 tr) {
	MeshTri o = r;
	V3 c;
	o.n = R * r.n;
	for (size_t j = 0; j < r.gauss.size(); j++) o.gauss[j] = R * r.gauss[j];
	for (int j = 0; j < 3; j++) {
		o.v[j] = tr + R * (r.v[j] - c) + c;
		o.ep1[j] = tr + R * (r.ep1[j] - c) + c;
		o.ep2[j] = tr + R * (r.ep2[j] - c) + c;
		o.eout[j] = R * r.eout[j];
	}
	triBox(o);
	return o;
}

// AABBTree::add (AABB_Tree.h:39-103) for a static leaf (no margin), surface
// area cost, with haveAncestorsRecalculateAABB and rebalanceAncestors (:357-424).
static void bvSet(BVNode& n, const double* box) {
	for (int i = 0; i < 6; i++) n.box[i] = box[i];
	n.cost = boxSurface(box);
}
static void bvSwap(Mesh& M, int n1, int n2) {  // Node::swpNodes (:306-319)
	int p1 = M.nodes[n1].parent, p2 = M.nodes[n2].parent;
	int* s1 = M.nodes[p1].left == n1 ? &M.nodes[p1].left : &M.nodes[p1].right;
	int* s2 = M.nodes[p2].left == n2 ? &M.nodes[p2].left : &M.nodes[p2].right;
	*s1 = n2;
	*s2 = n1;
	M.nodes[n1].parent = p2;
	M.nodes[n2].parent = p1;
}
static void bvAdd(Mesh& M, int tri) {
	BVNode leaf;
	leaf.leaf = tri;
	for (int i = 0; i < 6; i++) leaf.tbox[i] = M.tris[tri].aabb[i];
	double mb[6];
	for (int i = 0; i < 3; i++) {
		mb[i] = M.tris[tri].aabb[i] - 0.0;
		mb[3 + i] = M.tris[tri].aabb[3 + i] + 0.0;
	}
	bvSet(leaf, mb);
	int nl = (int)M.nodes.size();
	M.nodes.push_back(leaf);
	if (M.root < 0) {
		M.root = nl;
		return;
	}
	int best = -1;
	double best_cost = std::numeric_limits<double>::infinity();
	std::vector<std::pair<int, double>> cand{{M.root, 0.0}};
	while (!cand.empty()) {
		auto c = cand.back();
		cand.pop_back();
		if (c.second >= best_cost) continue;
		double cb[6];
		boxCombine(M.nodes[nl].box, M.nodes[c.first].box, cb);
		double direct = boxSurface(cb);
		double cost = c.second + direct;
		if (cost < best_cost) {
			best = c.first;
			best_cost = cost;
		}
		if (M.nodes[c.first].leaf < 0) {
			double inh = c.second + direct - M.nodes[c.first].cost;
			cand.push_back({M.nodes[c.first].left, inh});
			cand.push_back({M.nodes[c.first].right, inh});
		}
	}
	int np = (int)M.nodes.size();
	M.nodes.push_back(BVNode());
	if (best == M.root) {
		M.root = np;
	} else {
		int op = M.nodes[best].parent;
		if (M.nodes[op].left == best) M.nodes[op].left = np;
		else M.nodes[op].right = np;
		M.nodes[np].parent = op;
	}
	M.nodes[np].left = best;
	M.nodes[best].parent = np;
	M.nodes[np].right = nl;
	M.nodes[nl].parent = np;
	for (int cur = M.nodes[nl].parent; cur >= 0; cur = M.nodes[cur].parent) {
		double cb[6];
		boxCombine(M.nodes[M.nodes[cur].left].box, M.nodes[M.nodes[cur].right].box, cb);
		bvSet(M.nodes[cur], cb);
	}
	const double inf = std::numeric_limits<double>::infinity();
	int cur = M.nodes[nl].parent;
	while (cur >= 0 && M.nodes[cur].parent >= 0) {
		int par = M.nodes[cur].parent;
		int sib = M.nodes[par].left == cur ? M.nodes[par].right : M.nodes[par].left;
		if (sib >= 0) {
			double cl[6], cr[6], sl[6], sr[6], clc, crc, slc, src;
			const BVNode& C = M.nodes[cur];
			const BVNode& S = M.nodes[sib];
			boxCombine(M.nodes[C.right].box, S.box, sl);
			boxCombine(M.nodes[C.left].box, S.box, sr);
			slc = boxSurface(sl);
			src = boxSurface(sr);
			if (S.leaf >= 0) {
				clc = inf;
				crc = inf;
			} else {
				boxCombine(M.nodes[S.right].box, C.box, cl);
				boxCombine(M.nodes[S.left].box, C.box, cr);
				clc = boxSurface(cl);
				crc = boxSurface(cr);
			}
			double dcl = clc - S.cost, dcr = crc - S.cost, dsl = slc - C.cost, dsr = src - C.cost;
			double mn = std::min<double>(dcl, std::min<double>(dcr, std::min<double>(dsl, dsr)));
			if (mn >= 0) {
			} else if (mn == dcl) {
				int sl_node = M.nodes[sib].left;
				bvSwap(M, cur, sl_node);
				for (int i = 0; i < 6; i++) M.nodes[sib].box[i] = cl[i];
				M.nodes[sib].cost = clc;
			} else if (mn == dcr) {
				int sr_node = M.nodes[sib].right;
				bvSwap(M, cur, sr_node);
				for (int i = 0; i < 6; i++) M.nodes[sib].box[i] = cr[i];
				M.nodes[sib].cost = crc;
			} else if (mn == dsl) {
				int cl_node = M.nodes[cur].left;
				bvSwap(M, sib, cl_node);
				for (int i = 0; i < 6; i++) M.nodes[cur].box[i] = sl[i];
				M.nodes[cur].cost = slc;
			} else {
				int cr_node = M.nodes[cur].right;
				bvSwap(M, sib, cr_node);
				for (int i = 0; i < 6; i++) M.nodes[cur].box[i] = sr[i];
				M.nodes[cur].cost = src;
			}
		}
		cur = M.nodes[cur].parent;
	}
}

static bool aabbOverlap(const double* a, const double* b) {  // AABB.h:10-14
	return !(a[3] < b[0] || a[0] > b[3] || a[4] < b[1] || a[1] > b[4] || a[5] < b[2] || a[2] > b[5]);
}

// AABBTree::getCollisionCandidatesWith (AABB_Tree.h:226-247): depth first,
// right child popped first.
static std::vector<int> bvQuery(const Mesh& M, const double* target) {
	std::vector<int> out;
	if (M.root < 0) return out;
	std::vector<int> st{M.root};
	while (!st.empty()) {
		int n = st.back();
		st.pop_back();
		const BVNode& N = M.nodes[n];
		if (N.leaf >= 0) {
			if (aabbOverlap(N.tbox, target)) out.push_back(N.leaf);
		} else if (aabbOverlap(N.box, target)) {
			st.push_back(N.left);
			st.push_back(N.right);
		}
	}
	return out;
}

// RigidBody::updateGeometry -> Polyhedron::recomputeFromReference + gen_AABB
// (RigidBody.cpp:281-296, ConvexPrimitive.cpp:660-676, :715-729).
static void worldGeometry(Body& b) {
	M3 R = b.q.rot();
	if (b.type == 5) {  // StaticMeshGeometry::recomputeFromReference + genAABB (Geometry.cpp:849-868)
		Mesh& M = *b.mesh;
		M.tris.resize(M.ref.size());
		M.nodes.clear();
		M.root = -1;
		for (size_t i = 0; i < M.ref.size(); i++) {
			M.tris[i] = triTransform(M.ref[i], R, b.com);
			bvAdd(M, (int)i);
		}
		for (int i = 0; i < 6; i++) b.aabb[i] = M.tris[0].aabb[i];
		for (size_t i = 1; i < M.tris.size(); i++) {
			double o[6];
			boxCombine(b.aabb, M.tris[i].aabb, o);
			for (int k = 0; k < 6; k++) b.aabb[k] = o[k];
		}
		return;
	}
	if (b.type == 1) {  // Sphere::recomputeFromReference / gen_AABB (ConvexPrimitive.cpp:433-443)
		b.sph_c = b.com + R * b.sph_ref;
		V3 r(b.radius, b.radius, b.radius);
		V3 mn = b.sph_c - r, mx = b.sph_c + r;
		b.aabb[0] = mn.x; b.aabb[1] = mn.y; b.aabb[2] = mn.z;
		b.aabb[3] = mx.x; b.aabb[4] = mx.y; b.aabb[5] = mx.z;
		return;
	}
	if (b.type == 3) {  // Cylinder::recomputeFromReference / gen_AABB (ConvexPrimitive.cpp:272-298)
		b.cyl_c = R * b.cyl_c_ref + b.com;
		b.cyl_ax = R * b.cyl_ax_ref;
		for (int i = 0; i < 14; i++) b.cyl_gv[i] = R * b.cyl_gv_ref[i];
		for (int i = 0; i < 8; i++) {
			b.cyl_top[i] = R * b.cyl_top_ref[i] + b.com;
			b.cyl_bot[i] = b.cyl_top[i] - b.cyl_ax * b.cyl_h;
		}
		const double inf = std::numeric_limits<double>::infinity();
		double mn[3] = {inf, inf, inf}, mx[3] = {-inf, -inf, -inf};
		for (int i = 0; i < 16; i++) {  // point cloud order: top[0], bot[0], top[1], ...
			const V3& p = (i & 1) ? b.cyl_bot[i >> 1] : b.cyl_top[i >> 1];
			mn[0] = std::min<double>(mn[0], p.x); mn[1] = std::min<double>(mn[1], p.y); mn[2] = std::min<double>(mn[2], p.z);
			mx[0] = std::max<double>(mx[0], p.x); mx[1] = std::max<double>(mx[1], p.y); mx[2] = std::max<double>(mx[2], p.z);
		}
		for (int i = 0; i < 3; i++) { b.aabb[i] = mn[i]; b.aabb[3 + i] = mx[i]; }
		return;
	}
	Hull& h = b.hull;
	h.pts.resize(h.ref_pts.size());
	for (size_t i = 0; i < h.ref_pts.size(); i++) h.pts[i] = b.com + R * h.ref_pts[i];
	h.wnorm.resize(h.gnorm.size());
	for (size_t i = 0; i < h.gnorm.size(); i++) h.wnorm[i] = R * h.gnorm[i];
	const double inf = std::numeric_limits<double>::infinity();
	double mn[3] = {inf, inf, inf}, mx[3] = {-inf, -inf, -inf};
	for (const V3& p : h.pts) {
		mn[0] = std::min<double>(mn[0], p.x); mn[1] = std::min<double>(mn[1], p.y); mn[2] = std::min<double>(mn[2], p.z);
		mx[0] = std::max<double>(mx[0], p.x); mx[1] = std::max<double>(mx[1], p.y); mx[2] = std::max<double>(mx[2], p.z);
	}
	for (int i = 0; i < 3; i++) { b.aabb[i] = mn[i]; b.aabb[3 + i] = mx[i]; }
}

// RigidBody::updateInertiaTensor (RigidBody.cpp:270-279).
static void updateInertia(Body& b) {
	b.q = b.q.normalize();
	M3 r = b.q.rot(), rc = b.q.conj().rot();
	b.I = r * b.I_ref * rc;
	b.I_inv = r * b.I_ref_inv * rc;
}

// ---------------------------------------------------------------- narrowphase
struct Extrema {  // CollisionDetect.h:57-70
	int min_pid = -1, max_pid = -1;
	double min_val = std::numeric_limits<double>::infinity(), max_val = -std::numeric_limits<double>::infinity();
};
static Extrema findExtrema(const Hull& h, V3 axis) {  // CollisionDetect.cpp:627-644
	Extrema e;
	for (size_t i = 0; i < h.pts.size(); i++) {
		double val = h.pts[i].dot(axis);
		if (val < e.min_val) { e.min_pid = (int)i; e.min_val = val; }
		if (val > e.max_val) { e.max_pid = (int)i; e.max_val = val; }
	}
	return e;
}
struct CheckNorm {
	int a_pid, b_pid;
	V3 n;
	double pen;
};
static CheckNorm satCheck(const Extrema& a, const Extrema& b, V3 n) {  // CollisionDetect.cpp:705-715
	double f = a.max_val - b.min_val, r = b.max_val - a.min_val;
	if (f < r) return {a.max_pid, b.min_pid, n, f};
	return {a.min_pid, b.max_pid, -n, r};
}
static Extrema recentered(const Hull& h, int g) {  // CollisionDetect.cpp:622-625, :726
	double nv = h.wnorm[g].dot(h.pts[h.gref_idx[g]]);
	double diff = nv - h.gref_val[g];
	Extrema e;
	e.min_pid = h.gmin_pid[g]; e.max_pid = h.gmax_pid[g];
	e.min_val = h.gmin[g] + diff; e.max_val = h.gmax[g] + diff;
	return e;
}

enum Origin { FACE, EDGE, VERTEX, CYLINDER_BARREL };
struct P2 { double x, y; };
static P2 sub2(P2 a, P2 b) { return {a.x - b.x, a.y - b.y}; }
static double dot2(P2 a, P2 b) { double s = 0; s += a.x * b.x; s += a.y * b.y; return s; }
struct Area {
	std::vector<P2> ps;
	std::vector<int> ids;
	int surface;
	Origin origin;
};

// findContactArea (CollisionDetect.cpp:237-280) with projectFace/projectEdge (:215-235).
static Area contactArea(const Hull& h, V3 n, V3 p, int pid, V3 u, V3 w) {
	int best_f = -1;
	double best_ft = 1.0;
	for (int f = 0; f < (int)h.adj_faces[pid].size(); f++) {
		int fi = h.adj_faces[pid][f];
		double c = faceNormal(h, fi).dot(n);
		if (1 - c < best_ft) { best_ft = 1 - c; best_f = fi; }
	}
	if (best_ft <= COS_TOL) {
		Area a;
		a.surface = h.surf_id[best_f];
		a.origin = FACE;
		for (int vi : h.faces[best_f]) {
			V3 v = h.pts[vi];
			a.ps.push_back({v.dot(u), v.dot(w)});
			a.ids.push_back(vi);
		}
		return a;
	}
	int best_e = -1;
	double best_et = 1.0;
	for (int ei : h.adj_edges[pid]) {
		V3 e1 = h.pts[h.edges[ei].first], e2 = h.pts[h.edges[ei].second];
		double s = std::abs((e2 - e1).normalize().dot(n));
		if (s < best_et) { best_et = s; best_e = ei; }
	}
	if (best_et <= SIN_TOL) {
		V3 e1 = h.pts[h.edges[best_e].first], e2 = h.pts[h.edges[best_e].second];
		Area a;
		a.ps = {{e1.dot(u), e1.dot(w)}, {e2.dot(u), e2.dot(w)}};
		a.ids = {h.edges[best_e].first, h.edges[best_e].second};
		a.surface = -1;
		a.origin = EDGE;
		return a;
	}
	Area a;
	a.ps = {{p.dot(u), p.dot(w)}};
	a.ids = {pid};
	a.surface = -1;
	a.origin = VERTEX;
	return a;
}

static uint32_t edgeID(uint16_t a, uint16_t b) {  // CollisionDetect.cpp:9-21
	int mn, mx;
	if (a < b) { mn = a; mx = b; } else { mn = b; mx = a; }
	return (0x0000FFFF & mn) + (0xFFFF0000 & ((uint32_t)mx << 16));
}
static P2 inDir(P2 p1, P2 p2) {  // CollisionDetect.cpp:392-396, NVec.h:50-52
	P2 d{p1.y - p2.y, p2.x - p1.x};
	double s = 0; s += d.x * d.x; s += d.y * d.y;
	double k = 1.0 / std::sqrt(s);
	return {d.x * k, d.y * k};
}
struct CEP {
	P2 pos;
	int32_t src;
	uint64_t magic;
	bool skip;
};
static bool ccw(const Area& c) {  // CollisionDetect.cpp:398-409
	P2 in = inDir(c.ps[0], c.ps[1]);
	for (size_t i = 2; i < c.ps.size(); i++) {
		double d = dot2(sub2(c.ps[i], c.ps[0]), in);
		if (d > 0.0000000001) return true;
		else if (d < -0.0000000001) return false;
	}
	return false;
}
static std::vector<CEP> clipPoly(const Area& c, uint32_t other_sid, bool flip) {  // CollisionDetect.cpp:423-442
	std::vector<CEP> out;
	for (size_t i = 0; i < c.ps.size(); i++) {
		uint64_t m = 0;
		if (flip) {
			m |= 0x00000000FFFFFFFFull & other_sid;
			m |= 0xFFFFFFFF00000000ull & (uint64_t(c.ids[i]) << 32);
		} else {
			m |= 0x00000000FFFFFFFFull & (uint64_t)(int64_t)c.ids[i];
			m |= 0xFFFFFFFF00000000ull & (uint64_t(other_sid) << 32);
		}
		out.push_back({c.ps[i], c.ids[i], m, false});
	}
	if (c.ps.size() > 2 && !ccw(c)) std::reverse(out.begin(), out.end());
	return out;
}
static CEP intersect(CEP e1, CEP e2, P2 cn, P2 sp, int32_t ceid, bool flip) {  // CollisionDetect.cpp:444-467
	P2 d = sub2(e2.pos, e1.pos);
	double t = dot2(cn, sub2(sp, e1.pos)) / dot2(cn, d);
	P2 ip{e1.pos.x + d.x * t, e1.pos.y + d.y * t};
	bool e2p = dot2(d, cn) > 0;
	int32_t ps = e2p ? e1.src : e2.src;
	uint64_t m = 0;
	if (flip) {
		m |= 0x00000000FFFFFFFFull & (uint64_t)(int64_t)ceid;
		m |= 0xFFFFFFFF00000000ull & (uint64_t(edgeID((uint16_t)e1.src, (uint16_t)e2.src)) << 32);
	} else {
		m |= 0x00000000FFFFFFFFull & edgeID((uint16_t)e1.src, (uint16_t)e2.src);
		m |= 0xFFFFFFFF00000000ull & (uint64_t)((uint64_t)(int64_t)ceid << 32);
	}
	return {ip, ps, m, !e2p};
}
enum ClipR { BOTH, P1ON_P2C, P2ON_P1C, NEITHER, ONLY_P1, ONLY_P2 };
static ClipR checkSeg(double a, double b, double c) {  // CollisionDetect.cpp:473-480
	if (a >= c && b >= c) return NEITHER;
	if (a == c && b < c) return P1ON_P2C;
	if (a < c && b == c) return P2ON_P1C;
	if (a < c && b >= c) return ONLY_P1;
	if (a >= c && b < c) return ONLY_P2;
	return BOTH;
}
static std::vector<CEP> clipByEdge(const std::vector<CEP>& c, P2 side, P2 sp, int32_t eid, bool flip) {  // :482-562
	std::vector<CEP> out;
	double cv = dot2(sp, side);
	if (c.size() >= 3) {
		double p1v = 0, p2v = 0;
		CEP p1, p2;
		for (size_t i = 0; i < c.size(); i++) {
			if (i > 0) { p1v = p2v; p1 = p2; }
			else { p1 = c[0]; p1v = dot2(p1.pos, side); }
			p2 = c[(i + 1) % c.size()];
			p2v = dot2(p2.pos, side);
			ClipR r = checkSeg(p1v, p2v, cv);
			if (r == NEITHER || r == P1ON_P2C) out.push_back(p1);
			else if (r == ONLY_P2) { out.push_back(p1); out.push_back(intersect(p1, p2, side, sp, eid, flip)); }
			else if (r == ONLY_P1) out.push_back(intersect(p1, p2, side, sp, eid, flip));
		}
		return out;
	} else if (c.size() == 2) {
		double a = dot2(c[0].pos, side), b = dot2(c[1].pos, side);
		ClipR r = checkSeg(a, b, cv);
		if (r == NEITHER) return c;
		if (r == P1ON_P2C) return {c[0]};
		if (r == P2ON_P1C) return {c[1]};
		if (r == ONLY_P1) return {intersect(c[0], c[1], side, sp, eid, flip), c[1]};
		if (r == ONLY_P2) return {c[0], intersect(c[0], c[1], side, sp, eid, flip)};
		return out;
	} else {
		if (dot2(c[0].pos, side) >= cv) return c;
		return out;
	}
}
static std::vector<CEP> clipAll(std::vector<CEP> c1, const std::vector<CEP>& c2, bool flip) {  // :564-576
	for (size_t i = 0; i < c2.size(); i++) {
		CEP e1 = c2[i];
		if (e1.skip) continue;
		CEP e2 = c2[(i + 1) % c2.size()];
		P2 dir = inDir(e1.pos, e2.pos);
		int32_t eid = (int32_t)edgeID((uint16_t)e1.src, (uint16_t)e2.src);
		c1 = clipByEdge(c1, dir, e1.pos, eid, flip);
		if (c1.empty()) return c1;
	}
	return c1;
}
static std::vector<CEP> clipContacts(const Area& c1, const Area& c2) {  // :578-620
	std::vector<CEP> p1 = clipPoly(c1, (uint32_t)c2.surface, false);
	std::vector<CEP> p2 = clipPoly(c2, (uint32_t)c1.surface, true);
	if (p1.size() == 1) return p1;
	if (p2.size() == 1) return p2;
	if (p1.size() == 2 && p2.size() == 2) {
		P2 nn = inDir(p2[0].pos, p2[1].pos);
		return {intersect(p1[0], p1[1], nn, p2[0].pos, (int32_t)edgeID((uint16_t)p2[0].src, (uint16_t)p2[1].src), false)};
	}
	if (p1.size() == 2) return clipAll(p1, p2, false);
	if (p2.size() == 2) return clipAll(p2, p1, true);
	p1 = clipAll(p1, p2, false);
	return clipAll(p2, p1, true);
}

struct CP {
	V3 pos;
	double pen;
	uint64_t cid, bid;
	double restitution, mu_k, mu_s;
};
struct Manifold {
	std::vector<CP> pts;
	V3 n;
	double max_pen = -1;
};

// SAT_PolyPoly (CollisionDetect.cpp:717-835).
static Manifold satPolyPoly(const Body& A, uint64_t a_key, const Body& B, uint64_t b_key) {
	Manifold out;
	const Hull& a = A.hull;
	const Hull& b = B.hull;
	CheckNorm mp{-1, -1, V3(), std::numeric_limits<double>::infinity()};
	for (size_t g = 0; g < a.wnorm.size(); g++) {
		if (a.gred[g]) continue;
		CheckNorm x = satCheck(recentered(a, (int)g), findExtrema(b, a.wnorm[g]), a.wnorm[g]);
		if (x.pen < 0) { out.max_pen = -1; return out; }
		if (x.pen < mp.pen) mp = x;
	}
	for (size_t g = 0; g < b.wnorm.size(); g++) {
		if (b.gred[g]) continue;
		CheckNorm x = satCheck(findExtrema(a, b.wnorm[g]), recentered(b, (int)g), b.wnorm[g]);
		if (x.pen < 0) { out.max_pen = -1; return out; }
		if (x.pen < mp.pen) mp = x;
	}
	for (auto arc1 : a.arcs)
		for (auto arc2 : b.arcs) {
			V3 a1 = a.wnorm[arc1.first], a2 = a.wnorm[arc1.second];
			V3 b1 = -b.wnorm[arc2.first], b2 = -b.wnorm[arc2.second];
			V3 aa = a1 + a2;
			if (aa.dot(b1) + aa.dot(b2) <= 0) continue;
			V3 ap = a1.cross(a2), bp = b1.cross(b2);
			if (ap.dot(b1) * ap.dot(b2) > 0 || bp.dot(a1) * bp.dot(a2) > 0) continue;
			V3 n = ap.cross(bp);
			if (n.magSqrd() == 0) continue;
			n = n.normalize();
			if (aa.dot(n) < 0) n = n * -1.0;
			CheckNorm x = satCheck(findExtrema(a, n), findExtrema(b, n), n);
			if (x.pen < 0) { out.max_pen = -1; return out; }
			if (x.pen < mp.pen) mp = x;
		}
	out.n = mp.n;
	V3 n = mp.n;
	V3 amax = a.pts[mp.a_pid], bmax = b.pts[mp.b_pid];
	V3 u, w;
	n.perpBasis(&u, &w);
	Area ac = contactArea(a, n, amax, mp.a_pid, u, w);
	Area bc = contactArea(b, n * -1.0, bmax, mp.b_pid, u, w);
	std::vector<CEP> pool = clipContacts(ac, bc);
	double a_dot = amax.dot(n);
	V3 noff = n * a_dot;
	uint64_t cid = 0;
	cid |= 0x00000000FFFFFFFFull & a_key;
	cid |= 0xFFFFFFFF00000000ull & (b_key << 32);
	for (const CEP& p : pool) {
		CP c;
		c.pos = u * p.pos.x + w * p.pos.y + noff;
		c.pen = c.pos.dot(n) - a_dot + mp.pen;
		c.restitution = std::max<double>(A.restitution, B.restitution);
		c.mu_k = (A.mu_k + B.mu_k) / 2.0;
		c.mu_s = (A.mu_s + B.mu_s) / 2.0;
		c.cid = cid;
		c.bid = p.magic;
		out.pts.push_back(c);
	}
	out.max_pen = mp.pen;
	return out;
}

static Extrema sphereExtrema(const Body& s, V3 dir) {  // CollisionDetect.cpp:646-656
	Extrema e;
	double cv = dir.dot(s.sph_c);
	e.max_val = cv + s.radius;
	e.min_val = cv - s.radius;
	e.min_pid = -1;
	e.max_pid = -1;
	return e;
}

// detectSphereSphere (CollisionDetect.cpp:837-866).
static Manifold sphereSphere(const Body& A, uint64_t a_key, const Body& B, uint64_t b_key) {
	Manifold out;
	V3 diff = B.sph_c - A.sph_c;
	double d = diff.mag();
	out.max_pen = A.radius + B.radius - d;
	if (out.max_pen < 0) return out;
	out.n = diff.normalize();
	CP c;
	c.pos = A.sph_c + out.n * A.radius;
	c.pen = out.max_pen;
	c.cid = (0x00000000FFFFFFFFull & a_key) | (0xFFFFFFFF00000000ull & (b_key << 32));
	c.bid = 0;
	c.restitution = std::max<double>(A.restitution, B.restitution);
	c.mu_k = (A.mu_k + B.mu_k) / 2.0;
	c.mu_s = (A.mu_s + B.mu_s) / 2.0;
	out.pts.push_back(c);
	return out;
}

// SAT_PolySphere (CollisionDetect.cpp:1336-1411).
static Manifold satPolySphere(const Body& A, uint64_t a_key, const Body& B, uint64_t b_key) {
	Manifold out;
	out.max_pen = -1;
	const Hull& a = A.hull;
	CheckNorm mp{-1, -1, V3(), std::numeric_limits<double>::infinity()};
	uint32_t feat = 0;
	uint32_t sid = (uint32_t)a.pts.size();
	for (size_t g = 0; g < a.wnorm.size(); g++, sid++) {
		if (a.gred[g]) continue;
		CheckNorm x = satCheck(recentered(a, (int)g), sphereExtrema(B, a.wnorm[g]), a.wnorm[g]);
		if (x.pen < 0) { out.max_pen = -1; return out; }
		if (x.pen < mp.pen) { feat = sid; mp = x; }
	}
	for (size_t p = 0; p < a.pts.size(); p++) {
		V3 n = (a.pts[p] - B.sph_c).normalize();
		CheckNorm x = satCheck(findExtrema(a, n), sphereExtrema(B, n), n);
		if (x.pen < 0) { out.max_pen = -1; return out; }
		if (x.pen < mp.pen) { mp = x; feat = (uint32_t)p; }
	}
	for (auto& e : a.edges) {
		V3 e1 = a.pts[e.first], e2 = a.pts[e.second];
		V3 dir = (e2 - e1).normalize();
		V3 sample = e1 - B.sph_c;
		V3 n = (sample - dir * dir.dot(sample)).normalize();
		CheckNorm x = satCheck(findExtrema(a, n), sphereExtrema(B, n), n);
		if (x.pen < 0) { out.max_pen = -1; return out; }
		if (x.pen < mp.pen) { mp = x; feat = edgeID((uint16_t)e.first, (uint16_t)e.second); }
	}
	out.n = mp.n;
	CP c;
	c.pos = B.sph_c - out.n * B.radius;
	c.pen = mp.pen;
	c.restitution = std::max<double>(A.restitution, B.restitution);
	c.mu_k = (A.mu_k + B.mu_k) / 2.0;
	c.mu_s = (A.mu_s + B.mu_s) / 2.0;
	c.cid = (0x00000000FFFFFFFFull & a_key) | (0xFFFFFFFF00000000ull & (b_key << 32));
	c.bid = feat;
	out.pts.push_back(c);
	out.max_pen = mp.pen;
	return out;
}


// ---------------------------------------------------------------- cylinders
// Cylinder ids used in MagicIDs (ConvexPrimitive.h:174-179, 8-gon caps).
static const int CYL_TOP_PT = 0, CYL_BOT_PT = 8, CYL_TOP_SURF = 16, CYL_BOT_SURF = 17, CYL_TOP_EDGE = 18, CYL_BOT_EDGE = 19;

static V3 cylTopCenter(const Body& c) { return c.cyl_c + (0.5 * c.cyl_h) * c.cyl_ax; }  // ConvexPrimitive.h:182
static V3 cylBotCenter(const Body& c) { return c.cyl_c - (0.5 * c.cyl_h) * c.cyl_ax; }  // ConvexPrimitive.h:183

// Cylinder::getExtremaOfDisk (ConvexPrimitive.cpp:319-355).
static V3 diskExtrema(V3 dc, V3 dn, double r, V3 dir) {
	if (std::abs(dn.dot(dir)) > 0.99999999) return dc;
	if (std::abs(dn.dot(dir)) < 0.00000001) return dc + dir * r;
	V3 u, v, w;
	w = dir;
	dir.perpBasis(&u, &v);
	V3 un(dn.dot(u), dn.dot(v), dn.dot(w));
	double dwdu = -un.x / un.z;
	double dwdv = -un.y / un.z;
	double gm = std::sqrt(dwdu * dwdu + dwdv * dwdv);
	double gu = dwdu / gm;
	double gv = dwdv / gm;
	double uv2 = r * r / (1 + gm * gm);
	double dw = std::sqrt(r * r - uv2);
	double du = gu * std::sqrt(uv2);
	double dv = gv * std::sqrt(uv2);
	return dc + u * du + v * dv + w * dw;
}

// getCylinderExtrema (CollisionDetect.cpp:658-684).
static Extrema cylExtrema(const Body& c, V3 dir) {
	V3 tc = cylTopCenter(c);
	V3 tmax = diskExtrema(tc, c.cyl_ax, c.cyl_r, dir);
	V3 tmin = 2.0 * tc - tmax;
	V3 bmax = 2.0 * c.cyl_c - tmin;
	V3 bmin = 2.0 * c.cyl_c - tmax;
	double tmaxv = dir.dot(tmax), tminv = dir.dot(tmin), bmaxv = dir.dot(bmax), bminv = dir.dot(bmin);
	Extrema e;
	e.max_val = std::max<double>(tmaxv, bmaxv);
	e.min_val = std::min<double>(tminv, bminv);
	e.min_pid = -1;
	e.max_pid = -1;
	return e;
}

struct Seg { V3 p1, p2; };
// cylinderLengthwiseLineInDirection (CollisionDetect.cpp:347-352).
static Seg cylLine(V3 center, V3 ax, double r, double h, V3 n) {
	V3 barrel = (n - ax * ax.dot(n)).normalize();
	V3 p1 = center + barrel * r + (0.5 * h) * ax;
	V3 p2 = p1 - h * ax;
	return {p1, p2};
}

// findCylinderContactArea (CollisionDetect.cpp:354-374) with projectCylinderFace (:331-340).
static Area cylContactArea(const Body& c, V3 n, V3 u, V3 w) {
	V3 ax = c.cyl_ax;
	Area a;
	auto face = [&](const V3* fv, int surf, int off) {
		for (int i = 0; i < 8; i++) {
			a.ps.push_back({fv[i].dot(u), fv[i].dot(w)});
			a.ids.push_back(i + off);
		}
		a.surface = surf;
		a.origin = FACE;
	};
	if (1 - std::abs(ax.dot(n)) <= COS_TOL && n.dot(ax) > 0) face(c.cyl_top, CYL_TOP_SURF, CYL_TOP_PT);
	else if (1 - std::abs(ax.dot(n)) <= COS_TOL) face(c.cyl_bot, CYL_BOT_SURF, CYL_BOT_PT);
	else if (std::abs(ax.dot(n)) <= SIN_TOL) {
		Seg l = cylLine(c.cyl_c, ax, c.cyl_r, c.cyl_h, n);
		a.ps = {{l.p1.dot(u), l.p1.dot(w)}, {l.p2.dot(u), l.p2.dot(w)}};
		a.ids = {CYL_TOP_EDGE, CYL_BOT_EDGE};
		a.surface = -1;
		a.origin = CYLINDER_BARREL;
	} else if (n.dot(ax) > 0) {
		V3 p = diskExtrema(cylTopCenter(c), ax, c.cyl_r, n);
		a.ps = {{p.dot(u), p.dot(w)}};
		a.ids = {CYL_TOP_EDGE};
		a.surface = -1;
		a.origin = EDGE;
	} else {
		V3 p = diskExtrema(cylBotCenter(c), ax, c.cyl_r, n);
		a.ps = {{p.dot(u), p.dot(w)}};
		a.ids = {CYL_BOT_EDGE};
		a.surface = -1;
		a.origin = EDGE;
	}
	return a;
}

// Contact points from a projected pool (the lift shared by the SAT routines,
// e.g. CollisionDetect.cpp:1685-1708).
static void liftPool(Manifold& out, const std::vector<CEP>& pool, V3 n, V3 u, V3 w, V3 amax, double pen, const Body& A, uint64_t a_key,
                     const Body& B, uint64_t b_key) {
	double a_dot = amax.dot(n);
	V3 noff = n * a_dot;
	uint64_t cid = 0;
	cid |= 0x00000000FFFFFFFFull & a_key;
	cid |= 0xFFFFFFFF00000000ull & (b_key << 32);
	for (const CEP& p : pool) {
		CP c;
		c.pos = u * p.pos.x + w * p.pos.y + noff;
		c.pen = c.pos.dot(n) - a_dot + pen;
		c.restitution = std::max<double>(A.restitution, B.restitution);
		c.mu_k = (A.mu_k + B.mu_k) / 2.0;
		c.mu_s = (A.mu_s + B.mu_s) / 2.0;
		c.cid = cid;
		c.bid = p.magic;
		out.pts.push_back(c);
	}
	out.max_pen = pen;
}

static CEP poolPoint(P2 p, uint64_t magic) {
	CEP c{};
	c.pos = p;
	c.magic = magic;
	return c;
}

// Gauss-arc pair direction (the arc loop body of CollisionDetect.cpp:1617-1650 and
// :1021-1054); false when the pair is pruned.
static bool arcAxis(V3 a1, V3 a2, V3 b1, V3 b2, V3* out) {
	V3 aa = a1 + a2;
	if (aa.dot(b1) + aa.dot(b2) <= 0) return false;
	V3 ap = a1.cross(a2), bp = b1.cross(b2);
	if (ap.dot(b1) * ap.dot(b2) > 0 || bp.dot(a1) * bp.dot(a2) > 0) return false;
	V3 n = ap.cross(bp);
	if (n.magSqrd() == 0) return false;
	n = n.normalize();
	if (aa.dot(n) < 0) n = n * -1.0;
	*out = n;
	return true;
}

// Cylinder Gauss arcs as the reference builds them: {0, i} and {1, i} for
// i < 12 (ConvexPrimitive.cpp:239-245; the second index is the loop counter).
static void cylArc(int k, int* v1, int* v2) {
	*v1 = k & 1;
	*v2 = k >> 1;
}

// SAT_PolyCylinder (CollisionDetect.cpp:1563-1709).
static Manifold satPolyCylinder(const Body& A, uint64_t a_key, const Body& B, uint64_t b_key) {
	Manifold out;
	out.max_pen = -1;
	const Hull& a = A.hull;
	CheckNorm mp{-1, -1, V3(), std::numeric_limits<double>::infinity()};
	auto test = [&](const CheckNorm& x) {
		if (x.pen < 0) return true;
		if (x.pen < mp.pen) mp = x;
		return false;
	};
	for (size_t g = 0; g < a.wnorm.size(); g++) {
		if (a.gred[g]) continue;
		if (test(satCheck(recentered(a, (int)g), cylExtrema(B, a.wnorm[g]), a.wnorm[g]))) return Manifold();
	}
	V3 bax = B.cyl_ax;
	if (test(satCheck(findExtrema(a, bax), cylExtrema(B, bax), bax))) return Manifold();
	for (auto& e : a.edges) {
		V3 ed = a.pts[e.second] - a.pts[e.first];
		V3 dir = ed.cross(bax);
		if (dir.mag() < 0.00000000001) continue;
		V3 dn = dir.normalize();
		if (test(satCheck(findExtrema(a, dn), cylExtrema(B, dn), dn))) return Manifold();
	}
	for (const V3& p : a.pts) {
		V3 diff = p - B.cyl_c;
		V3 n = (diff - bax * bax.dot(diff)).normalize();
		if (test(satCheck(findExtrema(a, n), cylExtrema(B, n), n))) return Manifold();
	}
	for (auto arc1 : a.arcs)
		for (int k = 0; k < 24; k++) {
			int v1, v2;
			cylArc(k, &v1, &v2);
			V3 n;
			if (!arcAxis(a.wnorm[arc1.first], a.wnorm[arc1.second], -B.cyl_gv[v1], -B.cyl_gv[v2], &n)) continue;
			if (test(satCheck(findExtrema(a, n), cylExtrema(B, n), n))) return Manifold();
		}
	out.n = mp.n;
	V3 n = mp.n;
	V3 amax = a.pts[mp.a_pid];
	V3 u, w;
	n.perpBasis(&u, &w);
	Area ac = contactArea(a, n, amax, mp.a_pid, u, w);
	Area bc = cylContactArea(B, -n, u, w);
	std::vector<CEP> pool;
	if (bc.origin == EDGE) pool = {poolPoint(bc.ps[0], 0)};
	else pool = clipContacts(ac, bc);
	liftPool(out, pool, n, u, w, amax, mp.pen, A, a_key, B, b_key);
	return out;
}

// detectSphereCylinder (CollisionDetect.cpp:1712-1792).  The second disk
// centre repeats the top one (+h/2), as in the reference.
static Manifold sphereCylinder(const Body& A, uint64_t a_key, const Body& B, uint64_t b_key) {
	Manifold out;
	out.max_pen = -1;
	CheckNorm mp{-1, -1, V3(), std::numeric_limits<double>::infinity()};
	auto test = [&](V3 n) {
		CheckNorm x = satCheck(sphereExtrema(A, n), cylExtrema(B, n), n);
		if (x.pen < 0) return true;
		if (x.pen < mp.pen) mp = x;
		return false;
	};
	V3 ax = B.cyl_ax;
	if (test(ax)) return Manifold();
	V3 diff = B.cyl_c - A.sph_c;
	V3 barrel = (diff - ax * ax.dot(diff)).normalize();
	if (test(barrel)) return Manifold();
	V3 tdc = B.cyl_c + ax * 0.5 * B.cyl_h;
	V3 tcp = diskExtrema(tdc, ax, B.cyl_r, -diff.normalize());
	if (test((tcp - A.sph_c).normalize())) return Manifold();
	V3 bdc = B.cyl_c + ax * 0.5 * B.cyl_h;
	V3 bcp = diskExtrema(bdc, ax, B.cyl_r, -diff.normalize());
	if (test((bcp - A.sph_c).normalize())) return Manifold();
	out.n = mp.n;
	CP c;
	c.pos = A.sph_c + out.n * A.radius;
	c.pen = mp.pen;
	c.restitution = std::max<double>(A.restitution, B.restitution);
	c.mu_k = (A.mu_k + B.mu_k) / 2.0;
	c.mu_s = (A.mu_s + B.mu_s) / 2.0;
	c.cid = (0x00000000FFFFFFFFull & a_key) | (0xFFFFFFFF00000000ull & (b_key << 32));
	c.bid = 0;
	out.pts.push_back(c);
	out.max_pen = mp.pen;
	return out;
}

// checkCylinderEdgeApproxVsBarrel (CollisionDetect.cpp:938-956); true = separated.
static bool cylEdgeBarrel(const Body& A, const Body& B, V3 bax, const V3* edge, CheckNorm* mp) {
	for (int i = 0; i < 8; i++) {
		V3 ed = edge[(i + 1) % 8] - edge[i];
		V3 dir = ed.cross(bax);
		if (dir.mag() < 0.00000000001) continue;
		V3 dn = dir.normalize();
		CheckNorm x = satCheck(cylExtrema(A, dn), cylExtrema(B, dn), dn);
		if (x.pen < 0) return true;
		if (x.pen < mp->pen) *mp = x;
	}
	return false;
}

// detectCylinderCylinder (CollisionDetect.cpp:958-1137).
static Manifold cylCyl(const Body& A, uint64_t a_key, const Body& B, uint64_t b_key) {
	Manifold out;
	out.max_pen = -1;
	CheckNorm mp{-1, -1, V3(), std::numeric_limits<double>::infinity()};
	auto test = [&](V3 n) {
		CheckNorm x = satCheck(cylExtrema(A, n), cylExtrema(B, n), n);
		if (x.pen < 0) return true;
		if (x.pen < mp.pen) mp = x;
		return false;
	};
	V3 aax = A.cyl_ax, bax = B.cyl_ax;
	if (test(aax)) return Manifold();
	if (test(bax)) return Manifold();
	if (cylEdgeBarrel(A, B, bax, A.cyl_top, &mp)) return Manifold();
	if (cylEdgeBarrel(A, B, bax, A.cyl_bot, &mp)) return Manifold();
	if (cylEdgeBarrel(A, B, aax, B.cyl_top, &mp)) return Manifold();
	if (cylEdgeBarrel(A, B, aax, B.cyl_bot, &mp)) return Manifold();
	V3 bb = aax.cross(bax);
	if (bb.magSqrd() > 0.00000001) {
		bb = bb.normalize();
		if (test(bb)) return Manifold();
	}
	for (int k1 = 0; k1 < 24; k1++)
		for (int k2 = 0; k2 < 24; k2++) {
			int a1, a2, b1, b2;
			cylArc(k1, &a1, &a2);
			cylArc(k2, &b1, &b2);
			V3 n;
			if (!arcAxis(A.cyl_gv[a1], A.cyl_gv[a2], -B.cyl_gv[b1], -B.cyl_gv[b2], &n)) continue;
			if (test(n)) return Manifold();
		}
	out.n = mp.n;
	V3 n = mp.n;
	V3 u, w;
	n.perpBasis(&u, &w);
	Area ac = cylContactArea(A, n, u, w);
	Area bc = cylContactArea(B, -n, u, w);
	std::vector<CEP> pool;
	if (ac.origin == EDGE) pool = {poolPoint(ac.ps[0], 0)};
	else if (bc.origin == EDGE) pool = {poolPoint(bc.ps[0], 0)};
	else if (ac.origin == CYLINDER_BARREL && bc.origin == CYLINDER_BARREL && std::abs(aax.dot(bax)) > 0.995) {
		// parallel barrels (:1082-1104): overlap of the two lines along a's axis
		Seg la = cylLine(A.cyl_c, aax, A.cyl_r, A.cyl_h, n);
		Seg lb = cylLine(B.cyl_c, bax, B.cyl_r, B.cyl_h, -n);
		if (bax.dot(aax) < 0) std::swap(lb.p1, lb.p2);
		double amaxv = la.p1.dot(aax), aminv = la.p2.dot(aax), bmaxv = lb.p1.dot(aax), bminv = lb.p2.dot(aax);
		V3 imax = (amaxv > bmaxv) ? lb.p1 : la.p1;
		V3 imin = (aminv > bminv) ? la.p2 : lb.p2;
		pool = {poolPoint({u.dot(imax), w.dot(imax)}, 0x0), poolPoint({u.dot(imin), w.dot(imin)}, 0x1)};
	} else pool = clipContacts(ac, bc);
	V3 amax = aax.dot(n) > 0 ? diskExtrema(cylTopCenter(A), aax, A.cyl_r, n) : diskExtrema(cylBotCenter(A), aax, A.cyl_r, n);
	liftPool(out, pool, n, u, w, amax, mp.pen, A, a_key, B, b_key);
	return out;
}

static uint64_t swapHalves(uint64_t v) {  // swapOrder (CollisionDetect.cpp:28-40)
	uint64_t lo = v & 0x00000000FFFFFFFFull, hi = (v & 0xFFFFFFFF00000000ull) >> 32;
	return (0x00000000FFFFFFFFull & hi) + (0xFFFFFFFF00000000ull & (lo << 32));
}

// cull_manifold (CollisionDetect.cpp:2795-2830).
static Manifold cullManifold(const Manifold& m, int ns) {
	if (ns >= (int)m.pts.size()) return m;
	Manifold out;
	out.n = m.n;
	out.max_pen = 0;
	out.pts.resize(ns);
	std::vector<bool> avail(m.pts.size(), true);
	V3 u, w;
	m.n.perpBasis(&u, &w);
	for (int i = 0; i < ns; i++) {
		double vu = std::cos(2 * REF_PI * i / ns), vw = std::sin(2 * REF_PI * i / ns);
		V3 t = u * vu + w * vw;
		CP best;
		int bi = -1;
		double bv = -std::numeric_limits<double>::infinity();
		for (size_t j = 0; j < m.pts.size(); j++) {
			if (!avail[j]) continue;
			double v = m.pts[j].pos.dot(t);
			if (v > bv) { bv = v; best = m.pts[j]; bi = (int)j; }
		}
		out.pts[i] = best;
		out.max_pen = std::max<double>(out.max_pen, best.pen);
		avail[bi] = false;
	}
	return out;
}

// ---------------------------------------------------------------- convex vs triangle
static Extrema triExtrema(const MeshTri& t, V3 dir) {  // findTriangleExtrema (CollisionDetect.cpp:2004-2021)
	Extrema e;
	for (int i = 0; i < 3; i++) {
		double val = t.v[i].dot(dir);
		if (val < e.min_val) { e.min_pid = i; e.min_val = val; }
		if (val > e.max_val) { e.max_pid = i; e.max_val = val; }
	}
	return e;
}

// normalDirectionValid (CollisionDetect.cpp:2024-2060).
static bool normalValid(const MeshTri& t, V3 n) {
	const double EPS = 0.0001;
	switch (t.concave) {
	case 0:
	case 1:
		for (size_t i = 0; i < t.gauss.size(); i++) {
			V3 inner = t.gauss[i].cross(t.gauss[(i + 1) % t.gauss.size()]);
			if (n.dot(inner) < -EPS) return false;
		}
		break;
	case 2: {
		V3 an = t.gauss[0].cross(t.gauss[1]);
		if (std::abs(n.dot(an)) > EPS) return false;
		V3 up = an.cross(t.gauss[0]);
		if (n.dot(up) < -EPS) return false;
		V3 down = t.gauss[1].cross(an);
		if (n.dot(down) < -EPS) return false;
		break;
	}
	case 3:
		if (n.dot(t.n) < 1 - EPS) return false;
		break;
	}
	return true;
}

// findTriangleContactArea (CollisionDetect.cpp:282-329).
static Area triContactArea(const MeshTri& t, V3 n, V3 p, int pid, V3 u, V3 w) {
	Area a;
	if (1 - t.n.dot(n) <= COS_TOL) {
		for (int i = 0; i < 3; i++) {
			a.ps.push_back({t.v[i].dot(u), t.v[i].dot(w)});
			a.ids.push_back(i);
		}
		a.surface = t.id;
		a.origin = FACE;
		return a;
	}
	for (int i = 0; i < 3; i++) {
		if (i != pid && (i + 1) % 3 != pid) continue;
		if (std::abs((t.ep2[i] - t.ep1[i]).normalize().dot(n)) <= SIN_TOL) {
			a.ps = {{t.ep1[i].dot(u), t.ep1[i].dot(w)}, {t.ep2[i].dot(u), t.ep2[i].dot(w)}};
			a.ids = {t.eid[i], t.eid[i]};
			a.surface = -1;
			a.origin = EDGE;
			return a;
		}
	}
	a.ps = {{p.dot(u), p.dot(w)}};
	a.ids = {pid};
	a.surface = -1;
	a.origin = VERTEX;
	return a;
}

// Minimum-penetration bookkeeping shared by the triangle tests: the overall
// minimum and the minimum over Gauss-valid directions (normalDirectionValid
// of the reversed axis).
struct TriMin {
	CheckNorm all{-1, -1, V3(), std::numeric_limits<double>::infinity()};
	CheckNorm valid{-1, -1, V3(), std::numeric_limits<double>::infinity()};
	// false = separating axis
	bool test(const CheckNorm& x, const MeshTri& t, bool* improved = nullptr) {
		if (x.pen < 0) return false;
		if (improved) *improved = false;
		if (x.pen < all.pen) {
			all = x;
			if (improved) *improved = true;
		}
		if (x.pen < valid.pen && normalValid(t, -x.n)) valid = x;
		return true;
	}
	// the Gauss-valid minimum wins unless it is worse by more than the penalty
	CheckNorm pick(double penalty) const { return valid.pen < all.pen + penalty ? valid : all; }
};

// The six arcs of a triangle's Gauss map (face normal to each edge's outward
// direction and on to the reversed normal), CollisionDetect.cpp:2097-2103.
static void triArcs(const MeshTri& t, V3 g1[6], V3 g2[6]) {
	for (int e = 0; e < 3; e++) {
		g1[2 * e] = t.n;
		g2[2 * e] = t.eout[e];
		g1[2 * e + 1] = t.eout[e];
		g2[2 * e + 1] = -t.n;
	}
}

static void triPoint(CP& c, const Body& A, const MeshTri& t) {
	c.restitution = std::max<double>(A.restitution, t.rest);
	c.mu_k = (A.mu_k + t.mu_k) / 2.0;
	c.mu_s = (A.mu_s + t.mu_s) / 2.0;
}

// SAT_PolyTriangle (CollisionDetect.cpp:2062-2194).
static Manifold satPolyTriangle(const Body& A, uint64_t a_key, const MeshTri& t, double penalty) {
	Manifold out;
	out.max_pen = -1;
	const Hull& a = A.hull;
	TriMin m;
	if (!m.test(satCheck(findExtrema(a, t.n), triExtrema(t, t.n), t.n), t)) return Manifold();
	for (size_t g = 0; g < a.wnorm.size(); g++) {
		if (a.gred[g]) continue;
		if (!m.test(satCheck(recentered(a, (int)g), triExtrema(t, a.wnorm[g]), a.wnorm[g]), t)) return Manifold();
	}
	V3 g1[6], g2[6];
	triArcs(t, g1, g2);
	for (auto arc1 : a.arcs)
		for (int k = 0; k < 6; k++) {
			V3 n;
			if (!arcAxis(a.wnorm[arc1.first], a.wnorm[arc1.second], -g1[k], -g2[k], &n)) continue;
			if (!m.test(satCheck(findExtrema(a, n), triExtrema(t, n), n), t)) return Manifold();
		}
	CheckNorm raw = m.all, mp = m.pick(penalty);
	out.n = mp.n;
	V3 n = mp.n;
	V3 amax = a.pts[raw.a_pid], bmax = t.v[raw.b_pid];
	V3 u, w;
	n.perpBasis(&u, &w);
	Area ac = contactArea(a, raw.n, amax, raw.a_pid, u, w);
	Area bc = triContactArea(t, raw.n * -1.0, bmax, raw.b_pid, u, w);
	std::vector<CEP> pool = clipContacts(ac, bc);
	double a_dot = amax.dot(n);
	V3 noff = n * a_dot;
	uint64_t cid = (0x00000000FFFFFFFFull & a_key) | (0xFFFFFFFF00000000ull & ((uint64_t)t.id << 32));
	for (const CEP& p : pool) {
		CP c;
		c.pos = u * p.pos.x + w * p.pos.y + noff;
		c.pen = c.pos.dot(n) - a_dot + mp.pen;
		triPoint(c, A, t);
		c.cid = cid;
		c.bid = p.magic;
		out.pts.push_back(c);
	}
	out.max_pen = mp.pen;
	return out;
}

// SAT_SphereTriangle (CollisionDetect.cpp:2196-2276).
static Manifold satSphereTriangle(const Body& A, uint64_t a_key, const MeshTri& t, double penalty) {
	TriMin m;
	uint32_t feat = (uint32_t)-1;
	bool imp;
	if (!m.test(satCheck(sphereExtrema(A, t.n), triExtrema(t, t.n), t.n), t)) return Manifold();
	for (int i = 0; i < 3; i++) {
		V3 n = (t.v[i] - A.sph_c).normalize();
		if (!m.test(satCheck(sphereExtrema(A, n), triExtrema(t, n), n), t, &imp)) return Manifold();
		if (imp) feat = (uint32_t)i;
	}
	for (int e = 0; e < 3; e++) {
		V3 dir = (t.ep2[e] - t.ep1[e]).normalize();
		V3 sample = t.ep1[e] - A.sph_c;
		V3 n = (sample - dir * dir.dot(sample)).normalize();
		if (!m.test(satCheck(sphereExtrema(A, n), triExtrema(t, n), n), t, &imp)) return Manifold();
		if (imp) feat = (uint32_t)t.eid[e];
	}
	CheckNorm mp = m.pick(penalty);
	Manifold out;
	out.n = mp.n;
	CP c;
	c.pos = A.sph_c + mp.n * A.radius;
	c.pen = mp.pen;
	triPoint(c, A, t);
	c.cid = (0x00000000FFFFFFFFull & a_key) | (0xFFFFFFFF00000000ull & ((uint64_t)t.id << 32));
	c.bid = feat;
	out.pts.push_back(c);
	out.max_pen = mp.pen;
	return out;
}

// SAT_CylinderTriangle (CollisionDetect.cpp:2466-2645).
static Manifold satCylinderTriangle(const Body& A, uint64_t a_key, const MeshTri& t, double penalty) {
	TriMin m;
	V3 ax = A.cyl_ax;
	if (!m.test(satCheck(cylExtrema(A, ax), triExtrema(t, ax), ax), t)) return Manifold();
	if (!m.test(satCheck(cylExtrema(A, t.n), triExtrema(t, t.n), t.n), t)) return Manifold();
	for (int e = 0; e < 3; e++) {
		V3 dir = (t.ep2[e] - t.ep1[e]).cross(ax);
		if (dir.mag() < 0.00000000001) continue;
		V3 dn = dir.normalize();
		if (!m.test(satCheck(cylExtrema(A, dn), triExtrema(t, dn), dn), t)) return Manifold();
	}
	for (int i = 0; i < 3; i++) {
		V3 diff = t.v[i] - A.cyl_c;
		V3 n = (diff - ax * ax.dot(diff)).normalize();
		if (!m.test(satCheck(cylExtrema(A, n), triExtrema(t, n), n), t)) return Manifold();
	}
	V3 g1[6], g2[6];
	triArcs(t, g1, g2);
	for (int k1 = 0; k1 < 24; k1++)
		for (int k = 0; k < 6; k++) {
			int v1, v2;
			cylArc(k1, &v1, &v2);
			V3 n;
			if (!arcAxis(A.cyl_gv[v1], A.cyl_gv[v2], -g1[k], -g2[k], &n)) continue;
			if (!m.test(satCheck(cylExtrema(A, n), triExtrema(t, n), n), t)) return Manifold();
		}
	CheckNorm raw = m.all, mp = m.pick(penalty);
	Manifold out;
	out.n = mp.n;
	V3 n = mp.n;
	V3 amax = ax.dot(n) > 0 ? diskExtrema(cylTopCenter(A), ax, A.cyl_r, n) : diskExtrema(cylBotCenter(A), ax, A.cyl_r, n);
	V3 bmax = t.v[raw.b_pid];
	V3 u, w;
	n.perpBasis(&u, &w);
	Area ac = cylContactArea(A, raw.n, u, w);
	Area bc = triContactArea(t, -raw.n, bmax, raw.b_pid, u, w);
	std::vector<CEP> pool;
	if (ac.origin == EDGE) pool = {poolPoint(ac.ps[0], 0x0)};
	else pool = clipContacts(ac, bc);
	double a_dot = amax.dot(n);
	V3 noff = n * a_dot;
	uint64_t cid = (0x00000000FFFFFFFFull & a_key) | (0xFFFFFFFF00000000ull & ((uint64_t)t.id << 32));
	for (const CEP& p : pool) {
		CP c;
		c.pos = u * p.pos.x + w * p.pos.y + noff;
		c.pen = c.pos.dot(n) - a_dot + mp.pen;
		triPoint(c, A, t);
		c.cid = cid;
		c.bid = p.magic;
		out.pts.push_back(c);
	}
	out.max_pen = mp.pen;
	return out;
}

// SAT_{Poly,Sphere,Cylinder}Mesh (CollisionDetect.cpp:2647-2778) for a mesh
// held in world coordinates: the triangles whose boxes meet the convex body's
// box, in tree order, one manifold per colliding triangle.
static std::vector<Manifold> convexMesh(const Body& C, const Body& Mb) {
	std::vector<Manifold> out;
	const Mesh& M = *Mb.mesh;
	uint64_t key = (uint64_t)C.id * 16;
	for (int ti : bvQuery(M, C.aabb)) {
		const MeshTri& t = M.tris[ti];
		double pen = C.type == 0 ? 0 * std::min<double>(boxLongest(C.aabb), boxLongest(t.aabb))
		                         : 0.15 * std::min<double>(boxLongest(C.aabb), boxLongest(t.aabb));
		Manifold m;
		if (C.type == 0) m = satPolyTriangle(C, key, t, pen);
		else if (C.type == 1) m = satSphereTriangle(C, key, t, pen);
		else if (C.type == 3) m = satCylinderTriangle(C, key, t, pen);
		if (m.max_pen > 0 && !m.pts.empty()) out.push_back(m);
	}
	return out;
}

// merge_manifold (CollisionDetect.cpp:2780-2793).
static Manifold mergeManifold(const Manifold& a, const Manifold& b) {
	Manifold o;
	o.pts = a.pts;
	o.pts.insert(o.pts.end(), b.pts.begin(), b.pts.end());
	o.n = (a.n + b.n).normalize();
	o.max_pen = std::max<double>(a.max_pen, b.max_pen);
	return o;
}


// ---------------------------------------------------------------- constraints
struct ContactRow {  // ContactConstraint (ConstraintSolver.cpp:240-272)
	int a, b;
	NM<1, 6> ja, jb;
	NM<6, 1> ita, itb;
	double inv;  // impulse_to_value_inverse
	double target, ptarget, imp, pimp, bounce, cutoff;
};
struct FrictionRow {  // FrictionConstraint (ConstraintSolver.cpp:277-343)
	int a, b;
	NM<2, 6> ja, jb;
	NM<6, 2> ita, itb;
	NM<2, 2> inv;
	double target[2], imp[2], pimp[2];
	double mu;
	V3 u, w;
	bool static_ready;
};
struct Contact {
	int b1, b2;  // body indices
	uint64_t cid, bid;
	V3 lp1, lp2, ln1;
	double pen, bounce, mu, hardness, cutoff;
	int life;
	bool live;
	ContactRow c;
	FrictionRow f;
	uint64_t seq;  // insertion order within the edge
};

static double posCorrectCoeff(double s, double dt) { return std::min<double>(s * dt, 1.0 / dt); }  // PersistentConstraint.cpp:199

struct Joint {
	int kind;       // 1 distance, 3 ball-socket, 4 sliding hinge, 5 hinge, 6 weld, 7 slider, 8 cone limit, 9 twist limit,
	                // 10 motor, 11 piston drive, 12 piston slide limit, 13 spring
	int type;       // degree (rows): 1, 3, 4, 5, 6, 5, 1, 1, 1, 1, 1, 0
	int a, b;       // body indices (b1, b2 as created)
	V3 tp1, tp2;    // tracked points in the body frames (RigidBody.cpp:134-144)
	V3 ax1, ax2;    // hinge / slide axes in the body frames; distance: ax1.x = target distance
	double ps, rs;  // position / rotation correction strengths
	double imp[6] = {}, pimp[6] = {}, himp[6] = {};
	V3 u, w;        // basis of the current solver object (hinge, slider, sliding hinge)
	int sys = -1;   // holonomic system, -1 none
	NM<6, 6> ja, jb;
	NM<6, 6> ita, itb;
	NM<6, 6> inv;
	double tgt[6] = {}, ptgt[6] = {};
	// cone (8) and twist (9) limits: ConeRotationLimit / TwistLimit state
	// (PersistentConstraint.h:145-215); is_inactive starts true
	bool inactive = true;
	double cur_angle = 0;
	int above = 0;  // TwistLimitConstraint::lim_status == ABOVE_MAX
	// motor (10), piston drive (11) and its slide limit (12): Motor / Piston
	// (PersistentConstraint.h:218-290).  mode: 0 OFF, 1 constant torque or
	// force, 2 target velocity, 3 target position (PoweredConstraitMode)
	int mode = 0;
	double dmax = 0, dvel = 0, dpos = 0;  // max torque / force, target velocity, target position
	double jpos = 0;      // motor_angular_position / piston_position
	double prev_vel = 0;  // prev_velocity (the piston never writes it)
	int lstat = 0;        // limit status of the last build: 0 none, 1 below the minimum, 2 above the maximum
	double bound = 0;     // max_torque_impulse / max_impulse of the last build
	double tv = 0;        // the target velocity of the last build (real_target_velocity)
	double rest = 0;      // a spring's resting length
	double ccfm = std::numeric_limits<double>::quiet_NaN();  // setConstraintUseCustomCFM, NaN: the global CFM (CFM.h:5-12)
	bool active() const {
		if (kind == 13) return false;  // springs add no row (Spring::isSolverConstraint)
		if (kind == 10) return (mode != 0 && dmax != 0) || jpos < tp1.x || jpos > tp1.y;  // Motor::constraintIsActive (:110-112)
		if (kind == 11) return mode != 0 && dmax != 0;                                    // Piston::pistonIsActive (:195-197)
		return kind < 8 || !inactive;  // limits while active; slide limit while exceeded
	}
	// needsPosCorrect (ConstraintSolver.h:309, :350)
	bool pos_correct() const { return !(kind == 11 || (kind == 10 && lstat == 0)); }
};

struct HSys {
	std::vector<int> og, order;  // joint indices
	std::vector<int> table, vloc;
	int degree = 0, cap = 0;
	std::vector<double> buf, dbuf;
	bool pending = true;  // rows and factor to build at the next step start
};

struct World {
	orc_config cfg;
	double cutoff_vel;
	std::vector<Body> bodies;
	std::vector<Joint> joints;
	std::vector<HSys> systems;
	std::set<std::pair<int, int>> nocol;  // disallowCollision pairs (lo, hi body index)
	std::vector<int> node_order;          // island roots for steps given none (body indices)
	std::vector<Contact> cache;  // grouped by edge (b1 < b2 by id), edge order = insertion
	uint64_t seq = 0;
	std::vector<int64_t> last_rows, last_offs;
	std::vector<Contact> last_mid;
	// sleeping parameters (PhysicsEngine.h:218-224)
	int sleeping = 0;
	double sleep_delay = 0.5, vel_sleep_coeff = 0.1, ang_vel_eps = 0.1, accel_sleep_coeff = 0.044;
	int non_sleepy_tick_threshold = 3;
};

// RigidBody::wake (RigidBody.cpp:331-338).
static void wakeBody(Body& b) {
	if (!b.asleep) return;
	b.asleep = false;
	b.sleep_counter = 0;
	b.sleep_ticks = 0;
	b.history.clear();
}

static void applyCFMInv1(double itv, double cfm, double* inv) { *inv = 1.0 / (itv * (1 + cfm)); }

// Contact::updateSolverConstraints (PersistentConstraint.cpp:201-215) building
// ContactConstraint and FrictionConstraint.
static void buildRows(World& W, Contact& c, double sub_dt) {
	Body& A = W.bodies[c.b1];
	Body& B = W.bodies[c.b2];
	V3 p1 = A.q.rotate(c.lp1) + A.com;
	V3 p2 = B.q.rotate(c.lp2) + B.com;
	V3 n = A.q.rotate(c.ln1);
	double adj = (p2 - p1).dot(n) + c.pen;
	double cimp = c.c.imp * 1.0;
	double fimp[2] = {c.f.imp[0] * 1.0, c.f.imp[1] * 1.0};
	double pcc = posCorrectCoeff(c.hardness, sub_dt);
	double cfm = W.cfg.cfm;
	double ima = A.invMass(), imb = B.invMass();
	M3 ia = A.invTensor(), ib = B.invTensor();
	NM<6, 6> Ma = invMass6(ima, ia), Mb = invMass6(imb, ib);
	// contact row
	{
		ContactRow& r = c.c;
		r.a = c.b1; r.b = c.b2;
		V3 rA = p1 - A.com, rB = p1 - B.com;
		NM<3, 3> sa = crossNM(rA), sb = crossNM(rB);
		NM<1, 3> nd;
		nd.v[0][0] = n.x; nd.v[0][1] = n.y; nd.v[0][2] = n.z;
		NM<1, 3> ndn = neg(nd);
		NM<1, 3> ja_ang = mul(nd, sa), jb_ang = mul(ndn, sb);
		for (int i = 0; i < 3; i++) {
			r.ja.v[0][i] = ndn.v[0][i]; r.ja.v[0][3 + i] = ja_ang.v[0][i];
			r.jb.v[0][i] = nd.v[0][i]; r.jb.v[0][3 + i] = jb_ang.v[0][i];
		}
		r.ita = mul(Ma, tr(r.ja));
		r.itb = mul(Mb, tr(r.jb));
		NM<1, 1> t1 = mul(r.ja, r.ita), t2 = mul(r.jb, r.itb);
		double itv = t1.v[0][0] + t2.v[0][0];
		applyCFMInv1(itv, cfm, &r.inv);
		double slop = 0.05;
		double sp = std::max<double>(0.0, adj - slop);
		r.ptarget = sp * pcc;
		r.imp = cimp;
		r.pimp = 0;
		r.bounce = c.bounce;
		r.cutoff = c.cutoff;
	}
	// friction rows
	{
		FrictionRow& r = c.f;
		V3 su = r.u, sw = r.w;
		r.a = c.b1; r.b = c.b2;
		V3 rA = p1 - A.com, rB = p1 - B.com;
		n.perpBasis(&r.u, &r.w);
		V3 u = r.u, w = r.w;
		r.imp[0] = u.dot(su) * fimp[0] + u.dot(sw) * fimp[1];
		r.imp[1] = w.dot(su) * fimp[0] + w.dot(sw) * fimp[1];
		NM<1, 3> ud, wd;
		ud.v[0][0] = u.x; ud.v[0][1] = u.y; ud.v[0][2] = u.z;
		wd.v[0][0] = w.x; wd.v[0][1] = w.y; wd.v[0][2] = w.z;
		NM<3, 3> sa = crossNM(rA), sb = crossNM(rB);
		NM<1, 3> a_u = mul(neg(ud), sa), b_u = mul(ud, sb), a_w = mul(neg(wd), sa), b_w = mul(wd, sb);
		NM<1, 3> nu = neg(ud), nw = neg(wd);
		for (int i = 0; i < 3; i++) {
			r.ja.v[0][i] = ud.v[0][i]; r.ja.v[0][3 + i] = a_u.v[0][i];
			r.jb.v[0][i] = nu.v[0][i]; r.jb.v[0][3 + i] = b_u.v[0][i];
			r.ja.v[1][i] = wd.v[0][i]; r.ja.v[1][3 + i] = a_w.v[0][i];
			r.jb.v[1][i] = nw.v[0][i]; r.jb.v[1][3 + i] = b_w.v[0][i];
		}
		r.ita = mul(Ma, tr(r.ja));
		r.itb = mul(Mb, tr(r.jb));
		NM<2, 2> t1 = mul(r.ja, r.ita), t2 = mul(r.jb, r.itb), itv;
		for (int i = 0; i < 2; i++) for (int j = 0; j < 2; j++) itv.v[i][j] = t1.v[i][j] + t2.v[i][j];
		for (int i = 0; i < 2; i++) itv.v[i][i] *= (1 + cfm);
		// rowMajorOrderInverse<2> (NMat.cpp:11-21)
		double a = itv.v[0][0], bq = itv.v[0][1], cq = itv.v[1][0], d = itv.v[1][1];
		double det = a * d - bq * cq;
		double di = 1.0 / det;
		r.inv.v[0][0] = d * di; r.inv.v[0][1] = -bq * di;
		r.inv.v[1][0] = -cq * di; r.inv.v[1][1] = a * di;
		r.pimp[0] = r.pimp[1] = 0;
		r.mu = c.mu;
		r.static_ready = false;
	}
}

// ---------------------------------------------------------------- PGS (ConstraintSolver.cpp:80-213)
// ---------------------------------------------------------------- joints
// BallSocket / Hinge (PersistentConstraint.cpp:225-231, :274-285) with their
// solver constraints (ConstraintSolver.cpp:382-405, :468-550).  Degree n = 3
// (ball-socket) or 5 (hinge); rows use the first n of the fixed-size arrays.

// rowMajorOrderInverse<3> (NMat.cpp:24-36) and the generic Gauss-Jordan form
// (NMat.h:121-147) for n = 4..6.
static void inverseN(int n, const double* src, double* dst) {
	if (n == 1) { dst[0] = 1.0 / src[0]; return; }
	if (n == 3) {
		double a = src[0], b = src[1], c = src[2], d = src[3], e = src[4], f = src[5], g = src[6], h = src[7], i = src[8];
		double det = a * (e * i - f * h) - b * (d * i - f * g) + c * (d * h - e * g);
		double di = 1.0 / det;
		dst[0] = (e * i - f * h) * di; dst[1] = (c * h - b * i) * di; dst[2] = (b * f - c * e) * di;
		dst[3] = (f * g - d * i) * di; dst[4] = (a * i - c * g) * di; dst[5] = (c * d - a * f) * di;
		dst[6] = (d * h - e * g) * di; dst[7] = (b * g - a * h) * di; dst[8] = (a * e - b * d) * di;
		return;
	}
	double cp[36], t[36];
	for (int k = 0; k < n * n; k++) { cp[k] = src[k]; t[k] = 0; }
	for (int k = 0; k < n; k++) t[k * (n + 1)] = 1.0;
	for (int piv = 0; piv < n; piv++)
		for (int r = 0; r < n; r++) {
			if (piv == r) continue;
			double ratio = cp[n * r + piv] / cp[n * piv + piv];
			for (int c = 0; c < n; c++) {
				cp[n * r + c] -= ratio * cp[n * piv + c];
				t[n * r + c] -= ratio * t[n * piv + c];
			}
		}
	for (int piv = 0; piv < n; piv++)
		for (int c = 0; c < n; c++) t[n * piv + c] /= cp[n * piv + piv];
	for (int k = 0; k < n * n; k++) dst[k] = t[k];
}

// The solver constraint of a joint from the current body state.  source
// impulse / basis: the previous object's (warm start with coefficient 1.0,
// hinge angular impulses re-projected onto the new basis, ConstraintSolver.cpp:475-476).
// Quaternion::angleTo (Quaternion.h:35-39) and the rotation correction of
// the weld and slider rows (ConstraintSolver.cpp:688-696, :888-896): the
// rotation taking a's orientation to b's, as axis * angle, scaled by h.
static V3 rotCorrect(const Q& a, const Q& b, double h) {
	Q diff = b * a.conj();
	double dot = std::abs(a.r * b.r + a.i * b.i + a.j * b.j + a.k * b.k);
	double angle = dot > 0.99999999999 ? 0 : 2 * std::acos(std::abs(a.r * b.r + a.i * b.i + a.j * b.j + a.k * b.k));
	V3 axis;
	if (angle != 0) {  // getRotAxis (Quaternion.h:41-49)
		V3 ax(diff.i, diff.j, diff.k);
		axis = (diff.r < 0 ? -ax.normalize() : ax.normalize()) * h;
	}
	return angle * axis;
}
static NM<1, 3> rowNM(V3 v) {
	NM<1, 3> o;
	o.v[0][0] = v.x; o.v[0][1] = v.y; o.v[0][2] = v.z;
	return o;
}
template <int R, int C>
static NM<R, C> subNM(const NM<R, C>& a, const NM<R, C>& b) {  // NMat.h:64-72
	NM<R, C> o;
	for (int i = 0; i < R; i++) for (int k = 0; k < C; k++) o.v[i][k] = a.v[i][k] - b.v[i][k];
	return o;
}
// copyInto of a 1x3 row (NMat.h:36-42)
static void putRow(NM<6, 6>& m, int r, int c0, const NM<1, 3>& x) {
	for (int c = 0; c < 3; c++) m.v[r][c0 + c] = x.v[0][c];
}
static void putBlock(NM<6, 6>& m, int r0, int c0, const NM<3, 3>& x) {
	for (int r = 0; r < 3; r++) for (int c = 0; c < 3; c++) m.v[r0 + r][c0 + c] = x.v[r][c];
}
static NM<3, 3> idenNM() {  // idenMat<3> (NMat.h:110-119)
	NM<3, 3> o;
	for (int r = 0; r < 3; r++) o.v[r][r] = 1;
	return o;
}

// Solver rows of one joint from the current body state: the joint's
// updateSolverConstraints (PersistentConstraint.cpp:217-354, warm start
// coefficient 1) and its constraint's constructor (ConstraintSolver.cpp:
// 348-900).  reproject: a basis-carrying joint re-projects its stored
// impulses onto the new basis (false for holonomic rows rebuilt from a
// stored state).
// Taylor approximation of 1 - cos f used as the limits' position error
// (ConstraintSolver.cpp:425-428, :455-457).
static const double QUAT_PI = 3.1415926535;  // PI of Quaternion.h:4, used by the twist angle's unwrap

static double limitError(double f) {
	return -f * f * f * f * f * f * f * f / 40320.0 + f * f * f * f * f * f / 720.0 - f * f * f * f / 24.0 + f * f / 2.0;
}

// Motor::calculatePosition (PersistentConstraint.cpp:30-71): the angle
// predicted from the relative angular velocity, snapped to the measured angle
// of b2's comparison axis in b1's reference plane (the candidate nearest the
// prediction among the measured angle shifted by -2PI, 0, +2PI).
static double motorPosition(const Joint& j, const Body& A, const Body& B, V3 axis, V3 av1, V3 av2, double dt) {
	V3 u1, w1, u2, tmp;
	j.ax1.perpBasis(&u1, &w1);  // Motor ctor (:8-13)
	j.ax2.perpBasis(&u2, &tmp);
	V3 ru = A.q.rotate(u1), rw = A.q.rotate(w1), cmp = B.q.rotate(u2);
	double rel = axis.dot(av1) - axis.dot(av2);
	double pred = j.jpos + rel * dt;
	if (std::abs(cmp.dot(axis)) > 0.999) return pred;
	V3 cn = (cmp - axis * axis.dot(cmp)).normalize();
	double dv = cn.dot(ru);
	if (dv > 1.0) dv = 1.0;
	if (dv < -1.0) dv = -1.0;
	double relang = 2 * QUAT_PI - std::acos(dv);
	if (cn.dot(rw) < 0) relang = 2 * QUAT_PI - relang;
	double r = std::fmod(pred, 2 * QUAT_PI);
	double n2 = pred - r;
	const double cand[3] = {n2 + relang - 2 * QUAT_PI, n2 + relang, n2 + relang + 2 * QUAT_PI};
	double closest = std::numeric_limits<double>::infinity();
	for (double c : cand)
		if (std::abs(pred - c) < std::abs(pred - closest)) closest = c;
	return closest;
}

// getConstraintTargetVelocityValue of the motor (:77-108) and the piston
// (:153-185).  vrel: the relative velocity along the axis as each measures it;
// inv_sum: the inverse inertia (motor) or inverse mass (piston) sum along it.
// The reference tests abs(impulse < EPS), i.e. impulse < EPS.
static double driveTarget(const Joint& j, double vrel, double inv_sum, double imp, double pos) {
	switch (j.mode) {
	case 0: return 0;
	case 1: return (j.dmax > 0) ? std::numeric_limits<double>::infinity() : -std::numeric_limits<double>::infinity();
	case 2: return j.dvel;
	default: {
		double diff = pos - j.dpos;
		double dir = (diff > 0) ? -1 : 1;
		double vd = std::abs(vrel - j.prev_vel);
		const double EPS = 0.0000001;
		double inertia = (vd < EPS || imp < EPS) ? 1.0 / inv_sum : std::abs(imp) / vd;
		return dir * std::sqrt(2 * std::abs(diff) * j.dmax / inertia);
	}
	}
}

// Piston position b2_pos . axis - b1_pos . axis (Piston::calculatePosition, :149-151).
static double pistonPosition(V3 p1, V3 p2, V3 axis) { return p2.dot(axis) - p1.dot(axis); }

// A joint's state as the reference's constructors leave it, from the current
// bodies: limits inactive, a motor's angular position measured as the Motor
// constructor does (:17-22), a piston's position from its two attach points.
// The Piston constructor (:139-141) looks b2's point key up in b1's tracked
// points; the value is replaced at the first update and is seen only by a
// position query before the first step.
static void jointInitState(World& W, Joint& j) {
	for (int k = 0; k < 6; k++) j.imp[k] = j.pimp[k] = j.himp[k] = 0;
	j.u = V3();
	j.w = V3();
	j.inactive = true;
	j.cur_angle = 0;
	j.above = 0;
	j.lstat = 0;
	j.prev_vel = 0;
	j.jpos = 0;
	const Body& A = W.bodies[j.a];
	const Body& B = W.bodies[j.b];
	if (j.kind == 10) j.jpos = motorPosition(j, A, B, A.q.rotate(j.ax1), V3(), V3(), 0);
	if (j.kind == 13)  // addSpring (PhysicsEngine.cpp:978): a negative length takes the current distance
		j.rest = j.ax1.z >= 0 ? j.ax1.z : ((A.com + A.q.rotate(j.tp1)) - (B.com + B.q.rotate(j.tp2))).mag();
	if (j.kind == 11 || j.kind == 12) j.jpos = pistonPosition(A.com + A.q.rotate(j.tp1), B.com + B.q.rotate(j.tp2), A.q.rotate(j.ax1));
}

static void buildJoint(World& W, Joint& j, double sub_dt, bool reproject) {
	if (j.kind == 13) return;  // springs act through applySprings
	const Body& A = W.bodies[j.a];
	const Body& B = W.bodies[j.b];
	V3 pa = A.com + A.q.rotate(j.tp1);  // getTrackedP
	V3 pb = B.com + B.q.rotate(j.tp2);
	double cfm = std::isnan(j.ccfm) ? W.cfg.cfm : j.ccfm;  // cfm.getCFMValue(global_cfm)
	int n = j.type;
	double start[6];
	for (int k = 0; k < n; k++) start[k] = j.imp[k] * 1.0;  // warm_start_coefficient * impulse
	NM<6, 6> Ma = invMass6(A.invMass(), A.invTensor()), Mb = invMass6(B.invMass(), B.invTensor());
	NM<6, 6> ja, jb;
	double pcc = posCorrectCoeff(j.ps, sub_dt);
	double rcc = posCorrectCoeff(j.rs, sub_dt);
	if (j.kind == 8) {
		// ConeRotationLimit::updateSolverConstraints (PersistentConstraint.cpp:233-247):
		// inactive inside the cone (impulse cleared, rows left as they were)
		V3 d1 = A.q.rotate(j.ax1), d2 = B.q.rotate(j.ax2);
		double angle = std::acos(d1.dot(d2));
		j.inactive = angle < j.tp1.x;
		if (j.inactive) {
			j.imp[0] = 0.0;
			return;
		}
		// ConeLimitConstraint (ConstraintSolver.cpp:410-430), warm started from the stored impulse
		V3 h = d1.cross(d2).normalize();
		NM<1, 3> hd = rowNM(h);
		putRow(ja, 0, 3, hd);
		putRow(jb, 0, 3, neg(hd));
		j.ptgt[0] = limitError(angle - j.tp1.x) * j.rs;
	} else if (j.kind == 9) {
		// TwistLimit::updateSolverConstraints (PersistentConstraint.cpp:249-272)
		M3 r1 = A.q.rot(), r2 = B.q.rot();
		V3 wl, ul;
		j.ax1.perpBasis(&wl, &ul);  // TwistLimit ctor: getPerpendicularBasis(&b1_w_axis_local, &b1_u_axis_local)
		V3 t1 = r1 * j.ax1, u1 = r1 * ul, w1 = r1 * wl;
		V3 t2 = r2 * j.ax2;
		V3 r2w = r2 * wl;  // the reference rotates b1_w_axis_local here (PersistentConstraint.cpp:258)
		// getCurrentAngle (:115-132); its b2-axis alignment is computed but not used
		double prev = j.cur_angle;
		double cur = std::atan2(r2w.dot(u1), r2w.dot(w1));
		while (prev - cur > QUAT_PI) cur += 2 * QUAT_PI;
		while (prev - cur < -QUAT_PI) cur -= 2 * QUAT_PI;
		j.cur_angle = cur;
		double mn = j.tp1.x, mx = j.tp1.y;
		j.inactive = mn <= cur && cur <= mx;
		if (j.inactive) {
			j.imp[0] = 0.0;
			return;
		}
		// TwistLimitConstraint (ConstraintSolver.cpp:439-459)
		j.above = cur > mx ? 1 : 0;
		putRow(ja, 0, 3, neg(rowNM(t1)));
		putRow(jb, 0, 3, rowNM(t2));
		double f = j.above ? cur - mx : mn - cur;
		j.ptgt[0] = limitError(f) * j.rs;
	} else if (j.kind == 10) {
		// Motor::updateSolverConstraints (PersistentConstraint.cpp:320-329) ->
		// MotorConstraint (ConstraintSolver.cpp:555-584)
		V3 axis = A.q.rotate(j.ax1);
		V3 av1 = A.getAngVel(), av2 = B.getAngVel();
		j.jpos = motorPosition(j, A, B, axis, av1, av2, sub_dt);
		if (j.active()) {
			double inv_sum = axis.dot(A.invTensor() * axis) + axis.dot(B.invTensor() * axis);
			double target = 1.0 * driveTarget(j, axis.dot(av1) - axis.dot(av2), inv_sum, j.imp[0], j.jpos);
			j.bound = std::abs(j.dmax * sub_dt);
			double mn = j.tp1.x, mx = j.tp1.y;
			if (j.jpos > mx) {
				j.ptgt[0] = (mx - j.jpos) * rcc;
				j.tv = std::min<double>(0, target);
				j.lstat = 2;
			} else if (j.jpos < mn) {
				j.ptgt[0] = (mn - j.jpos) * rcc;
				j.tv = std::max<double>(0, target);
				j.lstat = 1;
			} else {
				j.ptgt[0] = 0;
				j.tv = target;
				j.lstat = 0;
			}
			NM<1, 3> ad = rowNM(axis);
			putRow(ja, 0, 3, ad);
			putRow(jb, 0, 3, neg(ad));
			j.prev_vel = axis.dot(av1) - axis.dot(av2);  // writePrevVel, after the build
		} else {
			j.prev_vel = axis.dot(av1) - axis.dot(av2);
			return;
		}
	} else if (j.kind == 11 || j.kind == 12) {
		// Piston::updateSolverConstraints (PersistentConstraint.cpp:300-318)
		V3 axis = A.q.rotate(j.ax1);
		j.jpos = pistonPosition(pa, pb, axis);
		if (j.kind == 12) {
			double mn = j.ax2.x, mx = j.ax2.y;
			j.inactive = !(j.jpos < mn || j.jpos > mx);  // slideLimitExceeded (:191-193)
			if (j.inactive) return;
			// SlideLimitConstraint (ConstraintSolver.cpp:736-766)
			if (j.jpos > mx) {
				j.ptgt[0] = (mx - j.jpos) * pcc;
				j.lstat = 2;
			} else {
				j.ptgt[0] = (mn - j.jpos) * pcc;
				j.lstat = 1;
			}
		} else {
			if (!j.active()) return;
			// PistonConstraint (ConstraintSolver.cpp:701-721)
			double inv_sum = A.invMass() + B.invMass();
			j.tv = driveTarget(j, axis.dot(A.getVel()) - axis.dot(B.getVel()), inv_sum, j.imp[0], j.jpos);
			j.bound = j.dmax * sub_dt;
		}
		NM<1, 3> sd = rowNM(axis);
		putRow(ja, 0, 0, neg(sd));
		putRow(ja, 0, 3, mul(neg(sd), crossNM(axis)));
		putRow(jb, 0, 0, sd);
	}
	for (int k = 0; k < n; k++) j.imp[k] = start[k];
	switch (j.kind) {
	case 8:
	case 9:
	case 10:
	case 11:
	case 12:
		break;
	case 1: {  // DistanceConstraint (ConstraintSolver.cpp:348-379), not in moving mode
		V3 rA = pa - A.com, rB = pb - B.com;
		V3 diff = pa - pb;
		V3 dd = diff.normalize();
		NM<1, 3> d = rowNM(dd);
		putRow(ja, 0, 0, d);
		putRow(ja, 0, 3, mul(neg(d), crossNM(rA)));
		putRow(jb, 0, 0, neg(d));
		putRow(jb, 0, 3, mul(d, crossNM(rB)));
		double target = j.ax1.x;
		double error = diff.magSqrd() < 0.00001 ? target : target - diff.normalize().dot(diff);
		j.ptgt[0] = pcc * error;
		break;
	}
	case 3: {  // BallSocketConstraint (:382-408)
		V3 rA = pa - A.com, rB = pb - B.com;
		putBlock(ja, 0, 0, idenNM());
		putBlock(ja, 0, 3, neg(crossNM(rA)));
		putBlock(jb, 0, 0, neg(idenNM()));
		putBlock(jb, 0, 3, crossNM(rB));
		V3 err = pb - pa;
		j.ptgt[0] = pcc * err.x; j.ptgt[1] = pcc * err.y; j.ptgt[2] = pcc * err.z;
		break;
	}
	case 5: {  // HingeConstraint (:468-553)
		V3 axa = A.q.rotate(j.ax1), axb = B.q.rotate(j.ax2);
		V3 su = j.u, sw = j.w;
		V3 rA = pa - A.com, rB = pa - B.com;  // both from the b1 point (ConstraintSolver.cpp:469)
		V3 u, w;
		axa.perpBasis(&u, &w);
		V3 nn = axb;
		if (reproject) {
			j.imp[3] = u.dot(su) * start[3] + u.dot(sw) * start[4];
			j.imp[4] = w.dot(su) * start[3] + w.dot(sw) * start[4];
		}
		j.u = u;
		j.w = w;
		putBlock(ja, 0, 0, idenNM());
		putBlock(ja, 0, 3, neg(crossNM(rA)));
		putBlock(jb, 0, 0, neg(idenNM()));
		putBlock(jb, 0, 3, crossNM(rB));
		putRow(ja, 3, 3, mul(rowNM(nn), crossNM(u)));
		putRow(jb, 3, 3, mul(rowNM(u), crossNM(nn)));
		putRow(ja, 4, 3, mul(rowNM(nn), crossNM(w)));
		putRow(jb, 4, 3, mul(rowNM(w), crossNM(nn)));
		V3 pc = (pb - pa) * pcc;
		j.ptgt[0] = pc.x; j.ptgt[1] = pc.y; j.ptgt[2] = pc.z;
		j.ptgt[3] = u.dot(nn) * rcc;
		j.ptgt[4] = w.dot(nn) * rcc;
		break;
	}
	case 7: {  // SliderConstraint (:611-697); Slider passes the position coefficient for both (PersistentConstraint.cpp:297)
		V3 axis = A.q.rotate(j.ax1).normalize();
		V3 u, w;
		axis.perpBasis(&u, &w);
		V3 su = j.u, sw = j.w;
		if (reproject) {
			j.imp[0] = u.dot(su) * start[0] + u.dot(sw) * start[1];
			j.imp[1] = w.dot(su) * start[0] + w.dot(sw) * start[1];
		}
		j.u = u;
		j.w = w;
		V3 rA = pa - A.com, rB = pb - B.com;
		V3 pe = pa - pb;
		NM<1, 3> ud = rowNM(u), wd = rowNM(w), ped = rowNM(pe);
		putRow(ja, 0, 0, ud);
		putRow(ja, 0, 3, subNM(mul(neg(ud), crossNM(rA)), mul(ped, crossNM(u))));
		putRow(jb, 0, 0, neg(ud));
		putRow(jb, 0, 3, mul(ud, crossNM(rB)));
		putRow(ja, 1, 0, wd);
		putRow(ja, 1, 3, subNM(mul(neg(wd), crossNM(rA)), mul(ped, crossNM(w))));
		putRow(jb, 1, 0, neg(wd));
		putRow(jb, 1, 3, mul(wd, crossNM(rB)));
		putBlock(ja, 2, 3, idenNM());
		putBlock(jb, 2, 3, neg(idenNM()));
		V3 rc = rotCorrect(A.q, B.q, pcc);
		j.ptgt[0] = -pe.dot(u) * pcc;
		j.ptgt[1] = -pe.dot(w) * pcc;
		j.ptgt[2] = rc.x; j.ptgt[3] = rc.y; j.ptgt[4] = rc.z;
		break;
	}
	case 4: {  // SlidingHingeConstraint (:781-857)
		V3 axa = A.q.rotate(j.ax1), axb = B.q.rotate(j.ax2);
		V3 u, w;
		axb.perpBasis(&u, &w);
		V3 nn = axa;
		V3 su = j.u, sw = j.w;
		if (reproject) {
			j.imp[2] = u.dot(su) * start[2] + u.dot(sw) * start[3];
			j.imp[3] = w.dot(su) * start[2] + w.dot(sw) * start[3];
		}
		j.u = u;
		j.w = w;
		V3 rA = pa - A.com, rB = pb - B.com;
		V3 pd = pa - pb;
		NM<1, 3> ud = rowNM(u), wd = rowNM(w), nd = rowNM(nn), pdd = rowNM(pd);
		putRow(ja, 0, 0, ud);
		putRow(ja, 0, 3, mul(neg(ud), crossNM(rA)));
		putRow(jb, 0, 0, neg(ud));
		putRow(jb, 0, 3, subNM(mul(ud, crossNM(rB)), mul(pdd, crossNM(u))));
		putRow(ja, 1, 0, wd);
		putRow(ja, 1, 3, mul(neg(wd), crossNM(rA)));
		putRow(jb, 1, 0, neg(wd));
		putRow(jb, 1, 3, subNM(mul(wd, crossNM(rB)), mul(pdd, crossNM(w))));
		putRow(ja, 2, 3, mul(neg(ud), crossNM(nn)));
		putRow(jb, 2, 3, mul(neg(nd), crossNM(u)));
		putRow(ja, 3, 3, mul(neg(wd), crossNM(nn)));
		putRow(jb, 3, 3, mul(neg(nd), crossNM(w)));
		j.ptgt[0] = -u.dot(pd) * pcc;
		j.ptgt[1] = -w.dot(pd) * pcc;
		j.ptgt[2] = -u.dot(nn) * rcc;
		j.ptgt[3] = -w.dot(nn) * rcc;
		break;
	}
	default: {  // 6: WeldConstraint (:861-897)
		V3 rA = pa - A.com, rB = pb - B.com;
		putBlock(ja, 0, 0, idenNM());
		putBlock(ja, 0, 3, neg(crossNM(rA)));
		putBlock(ja, 3, 3, idenNM());
		putBlock(jb, 0, 0, neg(idenNM()));
		putBlock(jb, 0, 3, crossNM(rB));
		putBlock(jb, 3, 3, neg(idenNM()));
		V3 pc = (pb - pa) * pcc;
		V3 rc = rotCorrect(A.q, B.q, rcc);
		j.ptgt[0] = pc.x; j.ptgt[1] = pc.y; j.ptgt[2] = pc.z;
		j.ptgt[3] = rc.x; j.ptgt[4] = rc.y; j.ptgt[5] = rc.z;
		break;
	}
	}
	// impulse_to_*_velocity = M^-1 J^T, impulse_to_value = J_a M_a^-1 J_a^T + J_b ...
	for (int r = 0; r < 6; r++)
		for (int c = 0; c < n; c++) {
			double s1 = 0, s2 = 0;
			for (int k = 0; k < 6; k++) { s1 += Ma.v[r][k] * ja.v[c][k]; s2 += Mb.v[r][k] * jb.v[c][k]; }
			j.ita.v[r][c] = s1;
			j.itb.v[r][c] = s2;
		}
	double itv[36];
	for (int r = 0; r < n; r++)
		for (int c = 0; c < n; c++) {
			double s1 = 0, s2 = 0;
			for (int k = 0; k < 6; k++) { s1 += ja.v[r][k] * j.ita.v[k][c]; s2 += jb.v[r][k] * j.itb.v[k][c]; }
			itv[r * n + c] = s1 + s2;
		}
	for (int r = 0; r < n; r++) itv[r * n + r] *= (1 + cfm);  // applyCFM (ConstraintSolver.cpp:919-926)
	double inv[36];
	inverseN(n, itv, inv);
	for (int r = 0; r < n; r++)
		for (int c = 0; c < n; c++) j.inv.v[r][c] = inv[r * n + c];
	j.ja = ja;
	j.jb = jb;
	for (int k = 0; k < n; k++) { j.pimp[k] = 0; j.himp[k] = 0; }
}

// Degreed constraint value J_a va + J_b vb (ConstraintSolver.h:86-88).
static void jointValue(const Joint& j, const double* va, const double* vb, double* out) {
	for (int r = 0; r < j.type; r++) {
		double x = 0, y = 0;
		for (int k = 0; k < 6; k++) { x += va[k] * j.ja.v[r][k]; y += vb[k] * j.jb.v[r][k]; }
		out[r] = x + y;
	}
}
// computeAndApplyVelocityChange (ConstraintSolver.h:71-74).
static void jointApply(const Joint& j, const double* imp, double* va, double* vb) {
	double da[6], db[6];
	for (int r = 0; r < 6; r++) {
		da[r] = 0; db[r] = 0;
		for (int k = 0; k < j.type; k++) { da[r] += imp[k] * j.ita.v[r][k]; db[r] += imp[k] * j.itb.v[r][k]; }
	}
	for (int r = 0; r < 6; r++) { va[r] += da[r]; vb[r] += db[r]; }
}

// ---------------------------------------------------------------- holonomic systems
// HolonomicSystem (HolonomicBlockSolver.cpp): min-fill elimination order,
// block-lower-triangular layout, block LDL factor and the solve d = A^-1 (t - c).
static const int HS_EMPTY = -1;

static bool jointsShareBody(const World& W, int c1, int c2) {
	const Joint& a = W.joints[c1];
	const Joint& b = W.joints[c2];
	return a.a == b.a || a.a == b.b || a.b == b.a || a.b == b.b;
}

// HolonomicSystem::HolonomicSystem (HolonomicBlockSolver.cpp:11-132).
static void hsysLayout(const World& W, HSys& s) {
	int n = (int)s.og.size();
	std::vector<char> edge(n * n, 0), empty(n * n, 0);
	std::vector<int> rem(n);
	for (int i = 0; i < n; i++) {
		rem[i] = i;
		for (int k = 0; k < n; k++) {
			if (jointsShareBody(W, s.og[i], s.og[k])) {
				empty[i * n + k] = empty[k * n + i] = 0;
				edge[i + k * n] = edge[k + i * n] = 1;
			} else {
				empty[i * n + k] = empty[k * n + i] = 1;
			}
		}
	}
	s.order.clear();
	while (!rem.empty()) {
		int best = -1, best_cnt = INT_MAX;
		std::vector<int> best_nb;
		for (int i = 0; i < (int)rem.size(); i++) {
			std::vector<int> nb;
			for (int k = 0; k < n; k++)
				if (k != rem[i] && edge[rem[i] + k * n]) nb.push_back(k);
			int cnt = 0;
			for (size_t p = 0; p < nb.size(); p++)
				for (size_t q = p + 1; q < nb.size(); q++)
					if (!edge[nb[p] + nb[q] * n]) cnt++;
			if (cnt < best_cnt) {
				best_cnt = cnt;
				best = i;
				best_nb = nb;
				if (cnt == 0) break;
			}
		}
		int el = rem[best];
		s.order.push_back(el);
		rem.erase(rem.begin() + best);
		for (int k = 0; k < n; k++) edge[el + k * n] = edge[k + el * n] = 0;
		for (size_t p = 0; p < best_nb.size(); p++)
			for (size_t q = p + 1; q < best_nb.size(); q++) {
				int a = best_nb[p], b = best_nb[q];
				empty[a * n + b] = empty[b * n + a] = 0;
				edge[a + b * n] = edge[b + a * n] = 1;
			}
	}
	// order holds og positions; convert to joint indices and build the table
	std::vector<int> pos = s.order;
	for (int& x : s.order) x = s.og[x];
	s.table.clear();
	s.vloc.clear();
	s.degree = 0;
	s.cap = 0;
	for (int col = 0; col < n; col++) {
		s.vloc.push_back(s.degree);
		s.degree += W.joints[s.order[col]].type;
		for (int row = col; row < n; row++) {
			if (empty[pos[row] * n + pos[col]]) s.table.push_back(HS_EMPTY);
			else {
				s.table.push_back(s.cap);
				s.cap += W.joints[s.order[row]].type * W.joints[s.order[col]].type;
			}
		}
	}
	s.buf.assign(s.cap, 0.0);
	s.dbuf.assign(6 * s.degree, 0.0);
}

static int hsLoc(const HSys& s, int row, int col) {  // getBlockBufferLocation (:603-614)
	int n = (int)s.order.size();
	return s.table[row - col + ((2 * n + 1) * col - col * col) / 2];
}
static int hsFirstBelow(const HSys& s, int col) {  // :616-622
	for (int row = col + 1; row < (int)s.order.size(); row++) {
		int l = hsLoc(s, row, col);
		if (l != HS_EMPTY) return l;
	}
	return HS_EMPTY;
}

// computeBlockInitialValue (:799-825, :714-744) for joints c1 (row) and c2 (col).
static void hsBlockInit(const World& W, double cfm, int c1, int c2, double* t) {
	const Joint& A = W.joints[c1];
	const Joint& B = W.joints[c2];
	int n1 = A.type, n2 = B.type;
	if (c1 == c2) {
		for (int r = 0; r < n1; r++)
			for (int c = 0; c < n1; c++) {
				double s1 = 0, s2 = 0;
				for (int k = 0; k < 6; k++) { s1 += A.ja.v[r][k] * A.ita.v[k][c]; s2 += A.jb.v[r][k] * A.itb.v[k][c]; }
				t[r * n1 + c] = s1 + s2;
			}
		for (int i = 0; i < n1; i++) t[i + n1 * i] *= (1 + cfm);
		return;
	}
	const NM<6, 6>* jac = nullptr;
	const NM<6, 6>* it = nullptr;
	if (A.a == B.a) { jac = &A.ja; it = &B.ita; }
	else if (A.a == B.b) { jac = &A.ja; it = &B.itb; }
	else if (A.b == B.a) { jac = &A.jb; it = &B.ita; }
	else if (A.b == B.b) { jac = &A.jb; it = &B.itb; }
	if (!jac) { for (int i = 0; i < n1 * n2; i++) t[i] = 0; return; }
	for (int i = 0; i < n1 * n2; i++) {  // matMult (:701-712)
		int row = i / n2, col = i % n2;
		t[i] = 0;
		for (int k = 0; k < 6; k++) t[i] += jac->v[row][k] * it->v[k][col];
	}
}

static double dotp(int n, const double* a, const double* b) {  // dotProd (:760-768)
	double o = 0;
	for (int i = 0; i < n; i++) o += a[i] * b[i];
	return o;
}

// HolonomicSystem::computeInverse (:459-543): block LDL in place.
static void hsFactor(const World& W, HSys& s, double cfm) {
	int nc = (int)s.order.size();
	for (int col = 0; col < nc; col++)
		for (int row = col; row < nc; row++) {
			int l = hsLoc(s, row, col);
			if (l == HS_EMPTY) continue;
			hsBlockInit(W, cfm, s.order[row], s.order[col], s.buf.data() + l);
		}
	for (int col = 0; col < nc; col++) {
		int dl = hsLoc(s, col, col);
		double* diag = s.buf.data() + dl;
		int bw = W.joints[s.order[col]].type;
		double tmp[36];
		inverseN(bw, diag, tmp);
		for (int k = 0; k < bw * bw; k++) diag[k] = tmp[k];
		if (col + 1 == nc) continue;
		int first = hsFirstBelow(s, col);
		for (int row = col + 1; row < nc; row++) {
			int l = hsLoc(s, row, col);
			if (l == HS_EMPTY) continue;
			double* tgt = s.dbuf.data() + (l - first);
			int bh = W.joints[s.order[row]].type;
			const double* src = s.buf.data() + l;
			for (int r = 0; r < bh; r++)
				for (int c = 0; c < bw; c++) tgt[c + bw * r] = dotp(bw, src + bw * r, diag + bw * c);
		}
		for (int ldc = col + 1; ldc < nc; ldc++)
			for (int ldr = ldc; ldr < nc; ldr++) {
				int l = hsLoc(s, ldr, col);
				if (l == HS_EMPTY || hsLoc(s, ldc, col) == HS_EMPTY) continue;
				const double* src = s.buf.data() + l;
				const double* lsrc = s.dbuf.data() + (hsLoc(s, ldc, col) - first);
				double* tgt = s.buf.data() + hsLoc(s, ldr, ldc);
				int w2 = W.joints[s.order[ldc]].type, h2 = W.joints[s.order[ldr]].type;
				for (int r = 0; r < h2; r++)
					for (int c = 0; c < w2; c++) tgt[r * w2 + c] -= dotp(bw, src + r * bw, lsrc + c * bw);
			}
		if (first == HS_EMPTY) continue;
		int total = hsLoc(s, col + 1, col + 1) - first;
		for (int k = 0; k < total; k++) s.buf[first + k] = s.dbuf[k];
	}
}

// multInverseWithVector (:223-296): v <- L^-T D^-1 L^-1 v.
static void hsSolve(const World& W, const HSys& s, std::vector<double>& v) {
	int nc = (int)s.order.size();
	int deg = s.degree;
	std::vector<double> eff(deg);
	for (int col = 0; col + 1 < nc; col++) {
		eff.assign(deg, 0.0);
		for (int row = col + 1; row < nc; row++) {
			int l = hsLoc(s, row, col);
			if (l == HS_EMPTY) continue;
			int bw = W.joints[s.order[col]].type, bh = W.joints[s.order[row]].type;
			const double* m = s.buf.data() + l;
			const double* x = v.data() + s.vloc[col];
			for (int i = 0; i < bh; i++) eff[s.vloc[row] + i] = dotp(bw, m + i * bw, x);
		}
		for (int i = 0; i < deg; i++) v[i] -= eff[i];
	}
	std::vector<double> d(deg);
	for (int col = 0; col < nc; col++) {
		int bw = W.joints[s.order[col]].type;
		const double* m = s.buf.data() + hsLoc(s, col, col);
		const double* x = v.data() + s.vloc[col];
		for (int i = 0; i < bw; i++) d[s.vloc[col] + i] = dotp(bw, m + i * bw, x);
	}
	v = d;
	std::vector<double> lt(6);
	for (int col = nc - 2; col >= 0; col--) {
		lt.assign(6, 0.0);
		int bw = W.joints[s.order[col]].type;
		for (int row = col + 1; row < nc; row++) {
			int l = hsLoc(s, row, col);
			if (l == HS_EMPTY) continue;
			int bh = W.joints[s.order[row]].type;
			const double* m = s.buf.data() + l;
			int si = s.vloc[row];
			for (int i = 0; i < bw; i++) {
				double dp = 0;
				for (int k = 0; k < bh; k++) dp += v[si + k] * m[i + bw * k];
				lt[i] += dp;
			}
		}
		for (int i = 0; i < bw; i++) v[s.vloc[col] + i] -= lt[i];
	}
}

struct DV {
	double v[6] = {0, 0, 0, 0, 0, 0};
	double p[6] = {0, 0, 0, 0, 0, 0};
	double l[6] = {0, 0, 0, 0, 0, 0};  // linearization-error detection (HolonomicBlockSolver.cpp:197-221)
};
static void addv(double* acc, const double* d) { for (int i = 0; i < 6; i++) acc[i] += d[i]; }

// Joint rows are (nullptr, joint index); contact rows (contact, 0 normal / 1 friction).
static void pgsIsland(World& W, const std::vector<std::pair<Contact*, int>>& rows, const std::vector<int>& systems, int hol_iters,
                      const std::vector<int>& island_bodies) {
	std::map<int, DV> dv;
	for (auto& r : rows) {
		if (r.first) { dv[r.first->b1]; dv[r.first->b2]; }
		else { dv[W.joints[r.second].a]; dv[W.joints[r.second].b]; }
	}
	auto bodyVel = [&](int i, double* v) {
		const Body& B = W.bodies[i];
		V3 l = B.getVel(), a = B.getAngVel();
		v[0] = l.x; v[1] = l.y; v[2] = l.z; v[3] = a.x; v[4] = a.y; v[5] = a.z;
	};
	// targets (ConstraintSolver.cpp:265-268 / ConstraintSolver.h:78)
	for (auto& r : rows) {
		Contact* c = r.first;
		if (!c) {
			Joint& j = W.joints[r.second];
			double va[6], vb[6], val[6];
			bodyVel(j.a, va);
			bodyVel(j.b, vb);
			jointValue(j, va, vb, val);
			if (j.kind == 10 || j.kind == 11) j.tgt[0] = j.tv - val[0];  // Motor / PistonConstraint::findTargetConstraintValue
			else
				for (int k = 0; k < j.type; k++) j.tgt[k] = -val[k];
			continue;
		}
		const Body& A = W.bodies[c->b1];
		const Body& B = W.bodies[c->b2];
		double va[6] = {A.getVel().x, A.getVel().y, A.getVel().z, A.getAngVel().x, A.getAngVel().y, A.getAngVel().z};
		double vb[6] = {B.getVel().x, B.getVel().y, B.getVel().z, B.getAngVel().x, B.getAngVel().y, B.getAngVel().z};
		if (r.second == 0) {
			double x[1], y[1];
			mulv(c->c.ja, va, x); mulv(c->c.jb, vb, y);
			double cur = x[0] + y[0];
			c->c.target = (cur < -c->c.cutoff) ? -(1 + c->c.bounce) * cur : -cur;
		} else {
			double x[2], y[2];
			mulv(c->f.ja, va, x); mulv(c->f.jb, vb, y);
			c->f.target[0] = -(x[0] + y[0]);
			c->f.target[1] = -(x[1] + y[1]);
		}
	}
	// warm start
	for (auto& r : rows) {
		Contact* c = r.first;
		if (!c) {
			Joint& j = W.joints[r.second];
			if (!isZero(j.imp, j.type)) jointApply(j, j.imp, dv[j.a].v, dv[j.b].v);
			continue;
		}
		DV& a = dv[c->b1];
		DV& b = dv[c->b2];
		if (r.second == 0) {
			if (isZero(&c->c.imp, 1)) continue;
			double da[6], db[6];
			mulv(c->c.ita, &c->c.imp, da); mulv(c->c.itb, &c->c.imp, db);
			addv(a.v, da); addv(b.v, db);
		} else {
			if (isZero(c->f.imp, 2)) continue;
			double da[6], db[6];
			mulv(c->f.ita, c->f.imp, da); mulv(c->f.itb, c->f.imp, db);
			addv(a.v, da); addv(b.v, db);
		}
	}
	// PGSConstraintStep<n> of a joint (ConstraintSolver.cpp:12-40); the limits project.
	auto jstep = [&](Joint& j, bool pseudo) {
		double* va = pseudo ? dv[j.a].p : dv[j.a].v;
		double* vb = pseudo ? dv[j.b].p : dv[j.b].v;
		double* acc = pseudo ? j.pimp : j.imp;
		const double* tgt = pseudo ? j.ptgt : j.tgt;
		int n = j.type;
		double cur[6], delta[6], m[6], diff[6];
		jointValue(j, va, vb, cur);
		for (int k = 0; k < n; k++) delta[k] = tgt[k] - cur[k];
		for (int i = 0; i < n; i++) {
			m[i] = 0;
			for (int k = 0; k < n; k++) m[i] += delta[k] * j.inv.v[i][k];
		}
		for (int k = 0; k < n; k++) diff[k] = (acc[k] + m[k]) - acc[k];
		if (j.kind >= 8) {  // projectValidImpulse (ConstraintSolver.cpp:432-435, :461-463, :591-605, :728-731, :768-775)
			double nw = acc[0] + m[0], pr;
			bool clamp = j.kind == 11 || (j.kind == 10 && j.lstat == 0);
			bool upper = (j.kind == 9 && j.above) || ((j.kind == 10 || j.kind == 12) && j.lstat == 2);
			if (clamp) pr = nw > 0 ? std::min<double>(nw, j.bound) : std::max<double>(nw, -j.bound);
			else pr = upper ? std::min<double>(nw, 0) : std::max<double>(nw, 0);
			diff[0] = pr - acc[0];
		}
		if (!isZero(diff, n)) {
			jointApply(j, diff, va, vb);
			for (int k = 0; k < n; k++) acc[k] += diff[k];
		}
	};
	auto step = [&](Contact* c, int kind, bool pseudo) {
		if (!c) { jstep(W.joints[kind], pseudo); return; }
		DV& A = dv[c->b1];
		DV& B = dv[c->b2];
		double* va = pseudo ? A.p : A.v;
		double* vb = pseudo ? B.p : B.v;
		if (kind == 0) {
			ContactRow& r = c->c;
			double* acc = pseudo ? &r.pimp : &r.imp;
			double tgt = pseudo ? r.ptarget : r.target;
			double x[1], y[1];
			mulv(r.ja, va, x); mulv(r.jb, vb, y);
			double cur = x[0] + y[0];
			double delta = tgt - cur;
			double m = 0; m += delta * r.inv;
			double nw = *acc + m;
			double diff = std::max<double>(nw, 0) - *acc;
			if (!isZero(&diff, 1)) {
				double da[6], db[6];
				mulv(r.ita, &diff, da); mulv(r.itb, &diff, db);
				addv(va, da); addv(vb, db);
				*acc += diff;
			}
		} else {
			FrictionRow& r = c->f;
			double x[2], y[2];
			mulv(r.ja, va, x); mulv(r.jb, vb, y);
			double cur[2] = {x[0] + y[0], x[1] + y[1]};
			double delta[2] = {r.target[0] - cur[0], r.target[1] - cur[1]};
			double m[2];
			mulv(r.inv, delta, m);
			double nw[2] = {r.imp[0] + m[0], r.imp[1] + m[1]};
			// FrictionConstraint::projectValidImpulse (ConstraintSolver.cpp:331-343)
			double maxm = r.mu * std::min<double>(std::numeric_limits<double>::infinity(), c->c.imp);
			double cm = std::sqrt(nw[0] * nw[0] + nw[1] * nw[1]);
			double pr[2];
			if (cm > maxm) {
				r.static_ready = false;
				double q = maxm / cm;
				pr[0] = nw[0] * q; pr[1] = nw[1] * q;
			} else {
				r.static_ready = true;
				pr[0] = nw[0]; pr[1] = nw[1];
			}
			double diff[2] = {pr[0] - r.imp[0], pr[1] - r.imp[1]};
			if (!isZero(diff, 2)) {
				double da[6], db[6];
				mulv(r.ita, diff, da); mulv(r.itb, diff, db);
				addv(va, da); addv(vb, db);
				r.imp[0] += diff[0]; r.imp[1] += diff[1];
			}
		}
	};
	for (int it = 0; it < W.cfg.vel_iters; it++)
		for (auto& r : rows) step(r.first, r.second, false);
	for (int it = 0; it < W.cfg.pos_iters; it++)
		for (auto& r : rows)
			if (r.first ? r.second == 0 : W.joints[r.second].pos_correct()) step(r.first, r.second, true);
	// block solve of the island's holonomic systems (ConstraintSolver.cpp:163-196)
	if (hol_iters > 0 && !systems.empty()) {
		auto apply = [&](HSys& hs, bool pseudo, double thr) {  // computeAndApplyImpulses (:298-355)
			std::vector<double> d(hs.degree);
			for (size_t i = 0; i < hs.order.size(); i++) {
				Joint& j = W.joints[hs.order[i]];
				double cur[6];
				jointValue(j, pseudo ? dv[j.a].p : dv[j.a].v, pseudo ? dv[j.b].p : dv[j.b].v, cur);
				const double* tgt = pseudo ? j.ptgt : j.tgt;
				for (int k = 0; k < j.type; k++) d[hs.vloc[i] + k] = tgt[k] - cur[k];
			}
			hsSolve(W, hs, d);
			if (thr > 0) {
				for (size_t i = 0; i < hs.order.size(); i++) {
					Joint& j = W.joints[hs.order[i]];
					const double* ch = d.data() + hs.vloc[i];
					if (isZero(ch, j.type)) continue;
					jointApply(j, ch, dv[j.a].l, dv[j.b].l);
				}
				for (size_t i = 0; i < hs.order.size(); i++) {
					Joint& j = W.joints[hs.order[i]];
					double r2 = 0;
					for (int k = 3; k < 6; k++) r2 += dv[j.a].l[k] * dv[j.a].l[k] + dv[j.b].l[k] * dv[j.b].l[k];
					if (r2 > thr * thr) return;  // discard the solution
				}
			}
			for (size_t i = 0; i < hs.order.size(); i++) {
				Joint& j = W.joints[hs.order[i]];
				const double* ch = d.data() + hs.vloc[i];
				if (isZero(ch, j.type)) continue;
				jointApply(j, ch, pseudo ? dv[j.a].p : dv[j.a].v, pseudo ? dv[j.b].p : dv[j.b].v);
				double* acc = pseudo ? j.pimp : j.himp;
				for (int k = 0; k < j.type; k++) acc[k] += ch[k];
			}
		};
		for (int it = 0; it < hol_iters; it++)
			for (int si : systems) {
				apply(W.systems[si], false, -1.0);
				apply(W.systems[si], true, (double)15.0f);  // excessive_linear_error_torque_threshold (PhysicsEngine.h:217)
			}
		// addScaledHolonomicImpulseToNextWarmStart with factor 0 (PhysicsEngine.cpp:373)
		for (auto& r : rows) {
			if (r.first) continue;
			Joint& j = W.joints[r.second];
			if (j.sys < 0) continue;
			for (int k = 0; k < j.type; k++) j.imp[k] += (double)0.0f * j.himp[k];
		}
	}
	// applyVelocityChange (PhysicsEngine.cpp:1113-1136)
	for (auto& kv : dv) {
		Body& b = W.bodies[kv.first];
		if (!b.dynamic()) continue;
		b.vel += V3(kv.second.v[0], kv.second.v[1], kv.second.v[2]);
		b.ang_vel += V3(kv.second.v[3], kv.second.v[4], kv.second.v[5]);
		b.pvel += V3(kv.second.p[0], kv.second.p[1], kv.second.p[2]);
		b.pang += V3(kv.second.p[3], kv.second.p[4], kv.second.p[5]);
	}
	// An asleep body whose new velocity exceeds wake_vel wakes its island
	// (wakeupIsland, :1285-1290); otherwise asleep bodies keep a zero
	// velocity.  The reference visits the bodies in the order of an
	// unordered_map keyed by address (ConstraintSolver.cpp:91, :199-212), so
	// an asleep body it visits before the waking one is zeroed first; this
	// restatement takes the order in which every waking body comes first.
	V3 g(W.cfg.gravity[0], W.cfg.gravity[1], W.cfg.gravity[2]);
	double wake_vel = 0.25 * W.vel_sleep_coeff * g.mag();
	bool wake = false;
	for (auto& kv : dv) {
		const Body& b = W.bodies[kv.first];
		if (b.getAsleep() && (b.vel.mag() > wake_vel || b.ang_vel.mag() > wake_vel)) wake = true;
	}
	if (wake) {
		for (int i : island_bodies) wakeBody(W.bodies[i]);
	} else {
		for (auto& kv : dv) {
			Body& b = W.bodies[kv.first];
			if (!b.getAsleep()) continue;
			b.vel = V3();
			b.ang_vel = V3();
		}
	}
}

// ---------------------------------------------------------------- integration
// RigidBody::rotateWhileApplyingGyroAccel (RigidBody.cpp:207-255), n_itr = 4.
static void gyroRotate(Body& b, double dt, int n_itr) {
	if (b.ang_vel.magSqrd() == 0) return;
	M3 T = b.I;
	Q o = b.q;
	V3 av = b.ang_vel;
	for (int i = 0; i < n_itr; i++) {
		V3 w = av;
		double ts = dt / n_itr;
		for (int s = 0; s < 2; s++) {
			V3 f = (w.cross(T * w)) * ts - T * (av - w);
			M3 jac = T + ts * (M3::cross_mat(w) * T - M3::cross_mat(T * w));
			if (std::abs(jac.det()) > 0.00000000001) w -= jac.inverse() * f;
		}
		av = w;
		o = Q(ts * av.mag(), av) * o;
		if (i + 1 < n_itr) {
			M3 r = o.rot(), rc = o.conj().rot();
			T = r * b.I_ref * rc;
		}
	}
	b.q = o;
	b.ang_vel = av;
}

// ---------------------------------------------------------------- step
static void refreshAll(World& W) {
	for (Body& b : W.bodies) {
		updateInertia(b);
		worldGeometry(b);
	}
}

static std::vector<std::pair<int, int>> broadphase(const World& W) {
	std::vector<std::pair<int, int>> v;
	for (size_t i = 0; i < W.bodies.size(); i++)
		for (size_t j = i + 1; j < W.bodies.size(); j++)
			if (aabbOverlap(W.bodies[i].aabb, W.bodies[j].aabb)) v.push_back({(int)i, (int)j});
	return v;
}

static uint64_t primKey(const Body& b) { return (uint64_t)b.id * 16; }

// detectCollision dispatch (CollisionDetect.cpp:57-152).
static Manifold narrowRaw(const World& W, int i, int j);
static Manifold narrow(const World& W, int i, int j) {
	Manifold m = narrowRaw(W, i, j);
	if (W.cfg.drop_nonfinite && !m.pts.empty()) {
		if (!std::isfinite(m.n.x) || !std::isfinite(m.n.y) || !std::isfinite(m.n.z)) return Manifold();
		std::vector<CP> keep;
		for (const CP& c : m.pts)
			if (std::isfinite(c.pos.x) && std::isfinite(c.pos.y) && std::isfinite(c.pos.z) && std::isfinite(c.pen)) keep.push_back(c);
		m.pts = keep;
	}
	return m;
}
static Manifold narrowRaw(const World& W, int i, int j) {
	const Body& A = W.bodies[i];
	const Body& B = W.bodies[j];
	// detectCollision dispatch (CollisionDetect.cpp:57-152): the swapped forms
	// negate the normal and swap both MagicID halves.
	Manifold m;
	auto swapped = [&](Manifold r) {
		r.n = -r.n;
		for (CP& c : r.pts) { c.cid = swapHalves(c.cid); c.bid = swapHalves(c.bid); }
		return r;
	};
	if (A.type == 0 && B.type == 0) m = satPolyPoly(A, primKey(A), B, primKey(B));
	else if (A.type == 0 && B.type == 1) m = satPolySphere(A, primKey(A), B, primKey(B));
	else if (A.type == 1 && B.type == 0) m = swapped(satPolySphere(B, primKey(B), A, primKey(A)));
	else if (A.type == 1 && B.type == 1) m = sphereSphere(A, primKey(A), B, primKey(B));
	else if (A.type == 0 && B.type == 3) m = satPolyCylinder(A, primKey(A), B, primKey(B));
	else if (A.type == 3 && B.type == 0) m = swapped(satPolyCylinder(B, primKey(B), A, primKey(A)));
	else if (A.type == 1 && B.type == 3) m = sphereCylinder(A, primKey(A), B, primKey(B));
	else if (A.type == 3 && B.type == 1) m = swapped(sphereCylinder(B, primKey(B), A, primKey(A)));
	else if (A.type == 3 && B.type == 3) m = cylCyl(A, primKey(A), B, primKey(B));
	if (!(m.max_pen > 0)) return Manifold();
	return cullManifold(m, 4);
}

// The manifolds of a body pair with a static mesh as the engine adds them
// (PhysicsEngine.cpp:223-277): per-triangle manifolds (normal flipped and
// MagicIDs swapped when the mesh is the first body, CollisionDetect.cpp:171-195),
// near-parallel ones merged into the first, each culled to 4 points.
static std::vector<Manifold> meshPair(const World& W, int i, int j) {
	const Body& A = W.bodies[i];
	const Body& B = W.bodies[j];
	std::vector<Manifold> ms;
	if (A.type == 5 && B.type == 5) return ms;
	if (B.type == 5) ms = convexMesh(A, B);
	else {
		ms = convexMesh(B, A);
		for (Manifold& m : ms) {
			m.n = -m.n;
			for (CP& c : m.pts) { c.cid = swapHalves(c.cid); c.bid = swapHalves(c.bid); }
		}
	}
	std::vector<Manifold> out;
	std::vector<bool> merged(ms.size(), false);
	for (size_t a = 0; a < ms.size(); a++) {
		if (merged[a]) continue;
		for (size_t b = a + 1; b < ms.size(); b++) {
			if (merged[b]) continue;
			double v = ms[a].n.dot(ms[b].n);
			if (1 - v < COS_TOL) {
				ms[a] = mergeManifold(ms[a], ms[b]);
				merged[b] = true;
			}
		}
		Manifold m = cullManifold(ms[a], 4);
		if (W.cfg.drop_nonfinite) {
			if (!std::isfinite(m.n.x) || !std::isfinite(m.n.y) || !std::isfinite(m.n.z)) continue;
			std::vector<CP> keep;
			for (const CP& c : m.pts)
				if (std::isfinite(c.pos.x) && std::isfinite(c.pos.y) && std::isfinite(c.pos.z) && std::isfinite(c.pen)) keep.push_back(c);
			m.pts = keep;
		}
		out.push_back(m);
	}
	return out;
}

// addContact (PhysicsEngine.cpp:1062-1111).
static void addContact(World& W, int i, int j, const CP& p, V3 n) {
	Body& A = W.bodies[i];
	Body& B = W.bodies[j];
	M3 ia = A.q.conj().rot(), ib = B.q.conj().rot();
	V3 nl = ia * n, pl1 = ia * (p.pos - A.com), pl2 = ib * (p.pos - B.com);
	for (Contact& c : W.cache) {
		if (c.b1 != i || c.b2 != j) continue;
		if (c.cid == p.cid && c.bid == p.bid) {
			double fr = c.f.static_ready ? p.mu_s : p.mu_k;
			c.mu = fr;
			c.pen = p.pen;
			c.ln1 = nl;
			c.lp1 = pl1;
			c.lp2 = pl2;
			c.life = W.cfg.contact_life;
			c.live = true;
			return;
		}
	}
	Contact c{};
	c.b1 = i; c.b2 = j;
	c.cid = p.cid; c.bid = p.bid;
	c.lp1 = pl1; c.lp2 = pl2; c.ln1 = nl;
	c.pen = p.pen; c.bounce = p.restitution; c.hardness = W.cfg.contact_hardness; c.mu = p.mu_k;
	c.cutoff = W.cutoff_vel;
	c.life = W.cfg.contact_life;
	c.live = true;
	c.c = ContactRow{};
	c.f = FrictionRow{};
	c.f.static_ready = false;
	c.seq = W.seq++;
	W.cache.push_back(c);
}

static void sortCache(World& W) {
	std::stable_sort(W.cache.begin(), W.cache.end(), [&](const Contact& a, const Contact& b) {
		int ai = W.bodies[a.b1].id, aj = W.bodies[a.b2].id, bi = W.bodies[b.b1].id, bj = W.bodies[b.b2].id;
		if (ai != bi) return ai < bi;
		if (aj != bj) return aj < bj;
		return a.seq < b.seq;
	});
}

// PhysicsEngine::bodySleepy (PhysicsEngine.cpp:1221-1254) over the movement
// history, oldest entry first.
static bool bodySleepy(const World& W, const Body& b) {
	if (b.sleep_disabled) return false;
	const int n = (int)b.history.size();
	const double dt = W.cfg.dt;
	if (n == 0 || 2 * n < (W.sleep_delay / dt)) return false;
	V3 av, aw;
	for (const auto& h : b.history) {
		av += h.first;
		aw += h.second;
	}
	av = V3(av.x / n, av.y / n, av.z / n);
	aw = V3(aw.x / n, aw.y / n, aw.z / n);
	V3 acc = (b.history.back().first - b.history.front().first) / (dt * (n - 1));
	V3 aacc = (b.history.back().second - b.history.front().second) / (dt * (n - 1));
	V3 g(W.cfg.gravity[0], W.cfg.gravity[1], W.cfg.gravity[2]);
	double vel_eps = W.vel_sleep_coeff * g.mag();
	double accel_eps = W.accel_sleep_coeff * g.mag();
	double ang_accel_eps = W.accel_sleep_coeff * 2.5;
	return av.mag() <= vel_eps && aw.mag() <= W.ang_vel_eps && acc.mag() <= accel_eps && aacc.mag() <= ang_accel_eps;
}

// PhysicsEngine::wakeupIsland (PhysicsEngine.cpp:1285-1290) through
// bfsVisitAll (:1263-1283) over the constraint graph as it stands (contact
// cache pairs and joint pairs): every dynamic body reachable through dynamic
// bodies from body i wakes.
static void wakeIsland(World& W, int i) {
	const int N = (int)W.bodies.size();
	std::vector<std::vector<int>> adj(N);
	for (const Contact& c : W.cache) {
		adj[c.b1].push_back(c.b2);
		adj[c.b2].push_back(c.b1);
	}
	for (const Joint& j : W.joints) {
		adj[j.a].push_back(j.b);
		adj[j.b].push_back(j.a);
	}
	std::vector<char> vis(N, 0);
	std::vector<int> st{i};
	while (!st.empty()) {
		int c = st.back();
		st.pop_back();
		if (vis[c]) continue;
		vis[c] = 1;
		if (W.bodies[c].dynamic()) wakeBody(W.bodies[c]);
		for (int o : adj[c])
			if (W.bodies[o].dynamic() && !vis[o]) st.push_back(o);
	}
}

// RigidBody::applyImpulse (RigidBody.cpp:60-66).  Its waking action only
// matters with sleeping on, which worlds with springs do not use here.
static void applyImpulse(Body& b, V3 imp, V3 pos) {
	V3 torque = (pos - b.com).cross(imp);
	b.vel += b.invMass() * imp;
	b.ang_vel += b.invTensor() * torque;
}

// Spring forces at a substep's start (PhysicsEngine.cpp:1313-1355), in the
// constraint graph's edge order: nodes in node-map order, a node's edges by
// the other body's id (insertNewEdge, :1984-1995), each edge once, an edge's
// springs in creation order.  getVelOfPoint (RigidBody.cpp:153-155) is taken
// at the other body's attach point, as the reference does.
static void applySprings(World& W, const std::vector<int>& roots, double dt) {
	std::vector<int> sp;
	for (size_t k = 0; k < W.joints.size(); k++)
		if (W.joints[k].kind == 13) sp.push_back((int)k);
	if (sp.empty()) return;
	std::vector<int> rank(W.bodies.size());
	for (size_t p = 0; p < roots.size(); p++) rank[roots[p]] = (int)p;
	auto key = [&](int k) {
		const Joint& j = W.joints[k];
		int first = rank[j.a] < rank[j.b] ? j.a : j.b, other = first == j.a ? j.b : j.a;
		return std::make_pair(rank[first], other);
	};
	std::stable_sort(sp.begin(), sp.end(), [&](int x, int y) { return key(x) < key(y); });
	for (int k : sp) {
		const Joint& j = W.joints[k];
		Body& B1 = W.bodies[j.a];
		Body& B2 = W.bodies[j.b];
		const double stiffness = j.ax1.x, damping = j.ax1.y;
		V3 p1 = B1.com + B1.q.rotate(j.tp1), p2 = B2.com + B2.q.rotate(j.tp2);
		V3 diff = p2 - p1;
		double distance = diff.mag();
		V3 dir = (distance != 0) ? (p2 - p1).normalize() : V3(0, -1, 0);
		double expand = stiffness * (distance - j.rest);
		applyImpulse(B1, dir * expand * dt, p1);
		applyImpulse(B2, -dir * expand * dt, p2);
		V3 rA = p1 - B1.com, rB = p2 - B2.com;
		V3 rdA = B1.invTensor() * dir.cross(rA), rdB = B2.invTensor() * dir.cross(rB);
		double sens = 1.0 / (B1.invMass() + B2.invMass() + rA.cross(rdA).dot(dir) + rB.cross(rdB).dot(dir));
		double velocity = (B2.vel + B2.ang_vel.cross(p1 - B2.com)).dot(dir) - (B1.vel + B1.ang_vel.cross(p2 - B1.com)).dot(dir);
		double dampen = damping * velocity;
		double max_dampen = velocity * sens / dt;
		if (velocity > 0 && dampen > max_dampen) dampen = max_dampen;
		else if (velocity < 0 && dampen < max_dampen) dampen = max_dampen;
		applyImpulse(B1, dir * dampen * dt, p1);
		applyImpulse(B2, -dir * dampen * dt, p2);
	}
}

// Returns 0, or -1 when a schedule's groups do not match the islands of a
// sleeping world.
static int step(World& W, const int64_t* node_order, const int64_t* sched_rows = nullptr, const int64_t* sched_offs = nullptr,
                int sched_n = 0) {
	const int N = (int)W.bodies.size();
	double dt = W.cfg.dt;
	V3 g(W.cfg.gravity[0], W.cfg.gravity[1], W.cfg.gravity[2]);
	// PhysicsEngine.cpp:65-87: waking actions, movement history and sleep
	// counters, gravity once per step
	const int hist_len = static_cast<int>(W.sleep_delay / dt);
	for (int i = 0; i < N; i++) {
		Body& b = W.bodies[i];
		if (b.wake_req) wakeIsland(W, i);
		b.wake_req = false;
		if (!(b.dynamic() && !b.getAsleep())) continue;
		b.history.push_back({b.vel, b.ang_vel});  // RigidBody::recordMovementState (RigidBody.cpp:347-352)
		while ((int)b.history.size() > hist_len) b.history.pop_front();
		if (bodySleepy(W, b)) {
			b.sleep_counter += dt;
		} else {
			b.sleep_ticks++;
			if (b.sleep_ticks >= W.non_sleepy_tick_threshold) {
				b.sleep_counter = 0.0;
				b.sleep_ticks = 0;
			}
		}
		b.vel += g * dt;
	}
	// broadphase + narrowphase (PhysicsEngine.cpp:102-335)
	auto pairs = broadphase(W);
	for (auto& pr : pairs) {
		const Body& A = W.bodies[pr.first];
		const Body& B = W.bodies[pr.second];
		// collision_possible (PhysicsEngine.cpp:174): a dynamic body, not both asleep
		if (!(A.dynamic() || B.dynamic()) || (A.getAsleep() && B.getAsleep())) continue;
		if (W.nocol.count({std::min(pr.first, pr.second), std::max(pr.first, pr.second)})) continue;  // collisionAllowed (:570-572)
		if (A.type == 5 || B.type == 5) {
			for (const Manifold& m : meshPair(W, pr.first, pr.second))
				for (const CP& p : m.pts) addContact(W, pr.first, pr.second, p, m.n);
			continue;
		}
		Manifold m = narrow(W, pr.first, pr.second);
		for (const CP& p : m.pts) addContact(W, pr.first, pr.second, p, m.n);
	}
	sortCache(W);
	W.last_mid = W.cache;
	const int hol = W.cfg.hol_iters;
	// holonomic systems whose rows and factor are not current (created, or
	// loaded from a state): rows from the current state with the stored
	// impulses as they are, then the factor (PhysicsEngine.cpp:1507-1522)
	if (hol > 0)
		for (HSys& hs : W.systems) {
			if (!hs.pending) continue;
			for (int ji : hs.og) buildJoint(W, W.joints[ji], dt / W.cfg.substeps, false);
			hsFactor(W, hs, W.cfg.cfm);
			hs.pending = false;
		}
	// island search (PhysicsEngine.cpp:1536-1613, :1263-1283)
	// edges: body pairs with contacts or joints; a node's edges sorted by the
	// other body's id (ConstraintGraphNode::insertNewEdge, PhysicsEngine.cpp:1984-1995);
	// an edge's rows are its joints (creation order), then its live contacts.
	struct Edge {
		int first = -1;            // first cache index of the pair, -1 none
		std::vector<int> joints;   // joint indices
	};
	std::map<std::pair<int, int>, Edge> edges;
	for (size_t k = 0; k < W.cache.size(); k++) {
		Edge& e = edges[{W.cache[k].b1, W.cache[k].b2}];
		if (e.first < 0) e.first = (int)k;
	}
	for (size_t ji = 0; ji < W.joints.size(); ji++) {
		int a = W.joints[ji].a, b = W.joints[ji].b;
		edges[{std::min(a, b), std::max(a, b)}].joints.push_back((int)ji);
	}
	std::vector<std::vector<std::pair<int, const Edge*>>> adj(N);
	for (auto& kv : edges) {
		adj[kv.first.first].push_back({kv.first.second, &kv.second});
		adj[kv.first.second].push_back({kv.first.first, &kv.second});
	}
	for (auto& a : adj)
		std::sort(a.begin(), a.end(), [&](const std::pair<int, const Edge*>& x, const std::pair<int, const Edge*>& y) {
			return W.bodies[x.first].id < W.bodies[y.first].id;
		});
	std::vector<std::vector<std::pair<Contact*, int>>> islands;
	std::vector<std::vector<int>> isl_sys, isl_bodies;
	std::vector<int> spring_roots;
	W.last_rows.clear();
	W.last_offs.assign(1, 0);
	{
		std::vector<char> visited(N, 0);
		std::vector<int> roots;
		if (node_order) for (int k = 0; k < N; k++) roots.push_back((int)node_order[k] - 1);
		else if ((int)W.node_order.size() == N) roots = W.node_order;
		else for (int k = 0; k < N; k++) roots.push_back(k);
		spring_roots = roots;
		for (int root : roots) {
			if (!W.bodies[root].dynamic() || visited[root]) continue;
			std::vector<std::pair<Contact*, int>> rows;
			std::vector<int> sys, bodies;
			std::vector<int64_t> lrows;
			bool all_ready = true;
			std::vector<int> st{root};
			while (!st.empty()) {
				int cur = st.back();
				st.pop_back();
				if (visited[cur]) continue;
				visited[cur] = 1;
				if (W.bodies[cur].dynamic()) {
					const Body& cb = W.bodies[cur];
					bodies.push_back(cur);
					for (auto& e : adj[cur]) {
						const Body& nb = W.bodies[e.first];
						if ((cb.getAsleep() && (nb.getAsleep() || nb.movement == 1)) || visited[e.first]) continue;
						int lo = std::min(cur, e.first), hi = std::max(cur, e.first);
						if (hol > 0)
							for (int ji : e.second->joints) {
								int si = W.joints[ji].sys;
								if (si >= 0 && std::find(sys.begin(), sys.end(), si) == sys.end()) sys.push_back(si);
							}
						for (int ji : e.second->joints) {
							// an inactive limit adds no solver row (PersistentConstraint.h:164)
							if (!W.joints[ji].active()) continue;
							rows.push_back({nullptr, ji});
							lrows.push_back(ji);
							lrows.push_back(2);
						}
						if (e.second->first < 0) continue;
						for (size_t k = e.second->first; k < W.cache.size() && W.cache[k].b1 == lo && W.cache[k].b2 == hi; k++) {
							if (!W.cache[k].live) continue;
							rows.push_back({&W.cache[k], 0});
							rows.push_back({&W.cache[k], 1});
							lrows.insert(lrows.end(), {(int64_t)k, 0, (int64_t)k, 1});
						}
					}
					// readyToSleep (PhysicsEngine.cpp:1256-1258)
					if (!(cb.sleep_counter >= W.sleep_delay)) all_ready = false;
				}
				for (auto& e : adj[cur])
					if (W.bodies[e.first].dynamic() && !visited[e.first]) st.push_back(e.first);
			}
			if (W.sleeping && all_ready) {
				// RigidBody::sleep (RigidBody.cpp:322-329)
				for (int i : bodies) {
					Body& b = W.bodies[i];
					b.asleep = true;
					b.vel = V3();
					b.ang_vel = V3();
				}
			} else if (!rows.empty()) {
				islands.push_back(rows);
				isl_sys.push_back(sys);
				isl_bodies.push_back(bodies);
				W.last_rows.insert(W.last_rows.end(), lrows.begin(), lrows.end());
				W.last_offs.push_back((int64_t)W.last_rows.size() / 2);
			}
		}
	}
	if (sched_rows) {
		// explicit sweep order over the post-narrowphase cache (one row list per
		// group); with sleeping the groups are the islands above (matched by
		// the island of each group's first dynamic body)
		if (W.sleeping && sched_n != (int)islands.size()) return -1;
		std::vector<int> isl_of(N, -1);
		for (size_t k = 0; k < isl_bodies.size(); k++)
			for (int i : isl_bodies[k]) isl_of[i] = (int)k;
		std::vector<std::vector<std::pair<Contact*, int>>> g_rows;
		std::vector<std::vector<int>> g_sys, g_bodies;
		W.last_rows.clear();
		W.last_offs.assign(1, 0);
		for (int g = 0; g < sched_n; g++) {
			std::vector<std::pair<Contact*, int>> rows;
			for (int64_t r = sched_offs[g]; r < sched_offs[g + 1]; r++) {
				if (sched_rows[2 * r + 1] == 2) rows.push_back({nullptr, (int)sched_rows[2 * r]});
				else rows.push_back({&W.cache[sched_rows[2 * r]], (int)sched_rows[2 * r + 1]});
			}
			if (!rows.empty()) {
				g_rows.push_back(rows);
				std::vector<int> sys;
				for (auto& r : rows)
					if (!r.first && W.joints[r.second].sys >= 0 &&
					    std::find(sys.begin(), sys.end(), W.joints[r.second].sys) == sys.end())
						sys.push_back(W.joints[r.second].sys);
				g_sys.push_back(sys);
				int first = -1;
				for (auto& r : rows) {
					int ends[2] = {r.first ? r.first->b1 : W.joints[r.second].a, r.first ? r.first->b2 : W.joints[r.second].b};
					for (int e : ends)
						if (first < 0 && W.bodies[e].dynamic()) first = e;
				}
				if (W.sleeping && (first < 0 || isl_of[first] < 0)) return -1;
				g_bodies.push_back(W.sleeping ? isl_bodies[isl_of[first]] : std::vector<int>());
			}
		}
		islands.swap(g_rows);
		isl_sys.swap(g_sys);
		isl_bodies.swap(g_bodies);
	}
	// substeps (PhysicsEngine.cpp:368-420)
	int S = W.cfg.substeps;
	for (int s = 0; s < S; s++) {
		double sdt = dt / S;
		bool first = s == 0, last = s == S - 1;
		applySprings(W, spring_roots, sdt);
		// maintainConstraintGraphApplyPoweredConstraints (PhysicsEngine.cpp:1378-1425)
		for (Contact& c : W.cache) {
			bool skip = !c.live;
			if (last) { c.life--; c.live = false; }
			if (skip) continue;
			buildRows(W, c, sdt);
		}
		// joints; holonomic ones are current at the first substep (PhysicsEngine.cpp:1396-1403)
		for (Joint& j : W.joints) {
			if (first && j.sys >= 0 && hol > 0) continue;
			buildJoint(W, j, sdt, true);
		}
		// removeExpiredContactConstraints (PhysicsEngine.cpp:1434-1446) at the
		// first substep, before the solve: Contact::isExpired
		// (PersistentConstraint.h:74) spares contacts of asleep bodies.  Island
		// pointers stay valid because expired contacts are never live; the
		// vector is compacted after this substep.
		std::vector<char> expired;
		if (first)
			for (const Contact& c : W.cache)
				expired.push_back(!W.bodies[c.b1].getAsleep() && !W.bodies[c.b2].getAsleep() && c.life < 0);
		for (size_t i = 0; i < islands.size(); i++) pgsIsland(W, islands[i], isl_sys[i], first ? hol : 0, isl_bodies[i]);
		// update_positions (PhysicsEngine.cpp:390-410)
		for (Body& b : W.bodies) {
			if (!b.dynamic() || b.getAsleep()) continue;
			b.prev_com = b.com;
			b.prev_q = b.q;
			b.com += (b.vel + b.pvel) * sdt;
			if (b.ang_vel.magSqrd() != 0) {
				gyroRotate(b, sdt, 4);
				V3 pr = b.pang;
				b.q = Q(sdt * pr.mag(), pr) * b.q;
			}
			b.pvel = V3();
			b.pang = V3();
			updateInertia(b);
		}
		if (first) {
			std::vector<Contact> keep;
			for (size_t k = 0; k < W.cache.size(); k++)
				if (!expired[k]) keep.push_back(W.cache[k]);
			// re-point island rows is unnecessary after this substep only if no later substep
			// uses them; rebuild the island vectors against the compacted cache instead.
			std::map<const Contact*, size_t> remap;
			size_t k = 0;
			for (size_t x = 0; x < W.cache.size(); x++) {
				if (!expired[x]) remap[&W.cache[x]] = k++;
			}
			for (auto& isl : islands)
				for (auto& r : isl)
					if (r.first) r.first = (Contact*)(uintptr_t)(remap[r.first] + 1);
			W.cache = keep;
			for (auto& isl : islands)
				for (auto& r : isl)
					if (r.first) r.first = &W.cache[(size_t)(uintptr_t)r.first - 1];
		}
	}
	// update_geometry (PhysicsEngine.cpp:424-438)
	for (Body& b : W.bodies) {
		if (!b.dynamic() || b.getAsleep()) continue;
		if (b.com != b.prev_com || b.q != b.prev_q) {
			b.q = b.q.normalize();
			worldGeometry(b);
		}
	}
	// calculateHolonomicSystemInversesAsync (PhysicsEngine.cpp:1615-1648): the
	// holonomic rows and factors the next step's first substep uses
	if (hol > 0)
		for (HSys& hs : W.systems) {
			for (int ji : hs.og) buildJoint(W, W.joints[ji], dt / W.cfg.substeps, true);
			hsFactor(W, hs, W.cfg.cfm);
		}
	return 0;
}

}  // namespace

struct orc_world : World {};

extern "C" {

orc_world* orc_create(const orc_config* cfg) {
	orc_world* w = new orc_world();
	w->cfg = *cfg;
	V3 g(cfg->gravity[0], cfg->gravity[1], cfg->gravity[2]);
	w->cutoff_vel = 2 * g.mag() * cfg->dt;  // PhysicsEngine.h:193
	return w;
}

void orc_destroy(orc_world* w) { delete w; }

int orc_add_hull_body(orc_world* w, int movement, const double* mp, const double* mat, int n_pts, const double* pts, int n_faces,
                      const int64_t* face_off, const int64_t* face_verts, const int64_t* surface_id, const int64_t* normal_calc,
                      const int64_t* normal_dir, const double* gn, const int64_t* gred, const int64_t* gint, const double* gval,
                      int n_arcs, const int64_t* arcs, int n_edges, const int64_t* edges) {
	Body b{};
	b.id = (int)w->bodies.size() + 1;
	b.movement = movement;
	b.mass = mp[0];
	for (int r = 0; r < 3; r++)
		for (int c = 0; c < 3; c++) {
			b.I_ref.v[r][c] = mp[1 + 3 * r + c];
			b.I_ref_inv.v[r][c] = mp[10 + 3 * r + c];
		}
	b.restitution = mat[1];
	b.mu_k = mat[2];
	b.mu_s = mat[3];
	Hull& h = b.hull;
	for (int i = 0; i < n_pts; i++) h.ref_pts.push_back(V3(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]));
	h.adj_faces.resize(n_pts);
	h.adj_edges.resize(n_pts);
	for (int f = 0; f < n_faces; f++) {
		std::vector<int> fv;
		for (int64_t k = face_off[f]; k < face_off[f + 1]; k++) fv.push_back((int)face_verts[k]);
		h.faces.push_back(fv);
		h.surf_id.push_back((int)surface_id[f]);
		h.ncalc.push_back((int)normal_calc[f]);
		h.ndir.push_back((int)normal_dir[f]);
		for (int v : fv) h.adj_faces[v].push_back(f);  // Polyhedron ctor, ConvexPrimitive.cpp:597-604
		h.gnorm.push_back(V3(gn[3 * f], gn[3 * f + 1], gn[3 * f + 2]));
		h.gred.push_back((int)gred[f]);
		h.gmin_pid.push_back((int)gint[3 * f]);
		h.gmax_pid.push_back((int)gint[3 * f + 1]);
		h.gref_idx.push_back((int)gint[3 * f + 2]);
		h.gmin.push_back(gval[3 * f]);
		h.gmax.push_back(gval[3 * f + 1]);
		h.gref_val.push_back(gval[3 * f + 2]);
	}
	for (int a = 0; a < n_arcs; a++) h.arcs.push_back({(int)arcs[2 * a], (int)arcs[2 * a + 1]});
	for (int e = 0; e < n_edges; e++) {
		h.edges.push_back({(int)edges[2 * e], (int)edges[2 * e + 1]});
		h.adj_edges[edges[2 * e]].push_back(e);
		h.adj_edges[edges[2 * e + 1]].push_back(e);
	}
	w->bodies.push_back(b);
	return (int)w->bodies.size() - 1;
}

int orc_add_sphere_body(orc_world* w, int movement, const double* mp, const double* mat, const double* center, double radius) {
	Body b{};
	b.id = (int)w->bodies.size() + 1;
	b.movement = movement;
	b.type = 1;
	b.mass = mp[0];
	for (int r = 0; r < 3; r++)
		for (int c = 0; c < 3; c++) {
			b.I_ref.v[r][c] = mp[1 + 3 * r + c];
			b.I_ref_inv.v[r][c] = mp[10 + 3 * r + c];
		}
	b.restitution = mat[1];
	b.mu_k = mat[2];
	b.mu_s = mat[3];
	b.sph_ref = V3(center[0], center[1], center[2]);
	b.radius = radius;
	w->bodies.push_back(b);
	return (int)w->bodies.size() - 1;
}

int orc_add_cylinder_body(orc_world* w, int movement, const double* mp, const double* mat, const double* params8, const double* gauss42,
                          const double* top24, const double* bot24) {
	Body b{};
	b.id = (int)w->bodies.size() + 1;
	b.movement = movement;
	b.type = 3;
	b.mass = mp[0];
	for (int r = 0; r < 3; r++)
		for (int c = 0; c < 3; c++) {
			b.I_ref.v[r][c] = mp[1 + 3 * r + c];
			b.I_ref_inv.v[r][c] = mp[10 + 3 * r + c];
		}
	b.restitution = mat[1];
	b.mu_k = mat[2];
	b.mu_s = mat[3];
	b.cyl_c_ref = V3(params8[0], params8[1], params8[2]);
	b.cyl_ax_ref = V3(params8[3], params8[4], params8[5]);
	b.cyl_r = params8[6];
	b.cyl_h = params8[7];
	for (int i = 0; i < 14; i++) b.cyl_gv_ref[i] = V3(gauss42[3 * i], gauss42[3 * i + 1], gauss42[3 * i + 2]);
	for (int i = 0; i < 8; i++) {
		b.cyl_top_ref[i] = V3(top24[3 * i], top24[3 * i + 1], top24[3 * i + 2]);
		b.cyl_bot_ref[i] = V3(bot24[3 * i], bot24[3 * i + 1], bot24[3 * i + 2]);
	}
	w->bodies.push_back(b);
	return (int)w->bodies.size() - 1;
}

int orc_add_joint(orc_world* w, int type, int b1, int b2, const double* g) {
	Joint j;
	static const int degree[14] = {0, 1, 0, 3, 4, 5, 6, 5, 1, 1, 1, 1, 1, 0};
	if (type < 1 || type > 13 || (degree[type] == 0 && type != 13)) return -1;
	j.kind = type;
	j.type = degree[type];
	j.a = b1;
	j.b = b2;
	j.tp1 = V3(g[0], g[1], g[2]);
	j.tp2 = V3(g[3], g[4], g[5]);
	j.ax1 = V3(g[6], g[7], g[8]);
	j.ax2 = V3(g[9], g[10], g[11]);
	j.ps = g[12];
	j.rs = g[13];
	w->joints.push_back(j);
	if (type != 13) w->nocol.insert({std::min(b1, b2), std::max(b1, b2)});  // addSpring allows the pair to collide
	return (int)w->joints.size() - 1;
}

int orc_add_mesh_body(orc_world* w, int n_pts, const double* pts, int n_tris, const int64_t* tri_idx, const double* tri_mat3) {
	Body b{};
	b.id = (int)w->bodies.size() + 1;
	b.movement = 1;
	b.mass = 0;
	b.type = 5;
	b.q = Q(1, 0, 0, 0);
	b.mesh = std::make_shared<Mesh>();
	meshBuild(*b.mesh, n_pts, pts, n_tris, tri_idx, tri_mat3);
	w->bodies.push_back(b);
	worldGeometry(w->bodies.back());
	return (int)w->bodies.size() - 1;
}

int orc_mesh_query(const orc_world* w, int body, const double* aabb6, int64_t* out, int cap) {
	std::vector<int> c = bvQuery(*w->bodies[body].mesh, aabb6);
	for (int i = 0; i < (int)c.size() && i < cap; i++) out[i] = c[i];
	return (int)c.size();
}

void orc_mesh_triangle(const orc_world* w, int body, int t, double* f40, int64_t* i8) {
	const MeshTri& r = w->bodies[body].mesh->tris[t];
	double* o = f40;
	auto put = [&](V3 v) { *o++ = v.x; *o++ = v.y; *o++ = v.z; };
	put(r.n);
	for (int k = 0; k < 3; k++) put(r.v[k]);
	for (int k = 0; k < 3; k++) put(r.eout[k]);
	for (int k = 0; k < 3; k++) put(k < (int)r.gauss.size() ? r.gauss[k] : V3());
	for (int k = 0; k < 6; k++) *o++ = r.aabb[k];
	*o++ = r.rest; *o++ = r.mu_k; *o++ = r.mu_s;
	*o++ = 0;
	for (int k = 0; k < 3; k++) { i8[k] = r.vid[k]; i8[3 + k] = r.eid[k]; }
	i8[6] = (int64_t)r.gauss.size();
	i8[7] = r.concave;
}

int orc_narrowphase_multi(const orc_world* w, int id_lo, int id_hi, int cap, double* heads, int64_t* counts, double* pts,
                          uint64_t* magic) {
	int i = id_lo - 1, j = id_hi - 1;
	std::vector<Manifold> ms;
	if (w->nocol.count({std::min(i, j), std::max(i, j)})) return 0;  // collisionAllowed (PhysicsEngine.cpp:570-572)
	if (!(w->bodies[i].dynamic() || w->bodies[j].dynamic())) return 0;
	if (w->bodies[i].type == 5 || w->bodies[j].type == 5) ms = meshPair(*w, i, j);
	else {
		Manifold m = narrow(*w, i, j);
		if (!m.pts.empty()) ms.push_back(m);
	}
	int off = 0;
	for (int k = 0; k < (int)ms.size() && k < cap; k++) {
		const Manifold& m = ms[k];
		heads[4 * k] = m.n.x; heads[4 * k + 1] = m.n.y; heads[4 * k + 2] = m.n.z; heads[4 * k + 3] = m.max_pen;
		counts[k] = (int64_t)m.pts.size();
		for (const CP& c : m.pts) {
			pts[4 * off] = c.pos.x; pts[4 * off + 1] = c.pos.y; pts[4 * off + 2] = c.pos.z; pts[4 * off + 3] = c.pen;
			magic[2 * off] = c.cid; magic[2 * off + 1] = c.bid;
			off++;
		}
	}
	return (int)ms.size();
}

void orc_set_joint_state(orc_world* w, const double* s) {
	for (size_t i = 0; i < w->joints.size(); i++) {
		Joint& j = w->joints[i];
		const double* r = s + 12 * i;
		for (int k = 0; k < 6; k++) j.imp[k] = r[k];
		if (j.kind == 8 || j.kind == 9) {
			// limits: is_inactive, the twist's tracked angle; an active twist
			// limit was built at that angle (lim_status)
			j.inactive = r[6] != 0;
			j.cur_angle = j.kind == 9 ? r[7] : 0.0;
			j.above = j.kind == 9 && !j.inactive && j.cur_angle > j.tp1.y;
			continue;
		}
		if (j.kind == 10 || j.kind == 11) {
			// motor / piston drive: position and prev_velocity; a motor row
			// built at that position has its limit status
			j.jpos = r[6];
			j.prev_vel = r[7];
			j.lstat = j.kind == 10 ? (j.jpos > j.tp1.y ? 2 : (j.jpos < j.tp1.x ? 1 : 0)) : 0;
			continue;
		}
		if (j.kind == 12) {  // slide limit: not exceeded, position
			j.inactive = r[6] != 0;
			j.jpos = r[7];
			j.lstat = j.jpos > j.ax2.y ? 2 : 1;
			continue;
		}
		if (j.kind == 13) {  // spring: resting length
			j.rest = r[6];
			continue;
		}
		j.u = V3(r[6], r[7], r[8]);
		j.w = V3(r[9], r[10], r[11]);
	}
	for (HSys& hs : w->systems) hs.pending = true;
}

void orc_get_joint_state(const orc_world* w, double* s) {
	for (size_t i = 0; i < w->joints.size(); i++) {
		const Joint& j = w->joints[i];
		double* r = s + 12 * i;
		for (int k = 0; k < 6; k++) r[k] = k < j.type ? j.imp[k] : 0.0;
		if (j.kind >= 8) {
			for (int k = 6; k < 12; k++) r[k] = 0.0;
			if (j.kind == 8 || j.kind == 9 || j.kind == 12) r[6] = j.inactive ? 1.0 : 0.0;
			if (j.kind == 9) r[7] = j.cur_angle;
			if (j.kind == 10 || j.kind == 11) {
				r[6] = j.jpos;
				r[7] = j.prev_vel;
			}
			if (j.kind == 12) r[7] = j.jpos;
			if (j.kind == 13) r[6] = j.rest;
			continue;
		}
		r[6] = j.u.x; r[7] = j.u.y; r[8] = j.u.z;
		r[9] = j.w.x; r[10] = j.w.y; r[11] = j.w.z;
	}
}

void orc_set_joint_cfm(orc_world* w, int k, double cfm) { w->joints[k].ccfm = cfm; }

void orc_set_node_order(orc_world* w, int n, const int64_t* ids) {
	w->node_order.clear();
	for (int k = 0; k < n; k++) w->node_order.push_back((int)ids[k] - 1);
}

void orc_init_joint_state(orc_world* w) {
	for (Joint& j : w->joints) jointInitState(*w, j);
	for (HSys& hs : w->systems) hs.pending = true;
}

int orc_set_drive(orc_world* w, int k, int mode, double max, double target) {
	if (k < 0 || k >= (int)w->joints.size() || mode < 0 || mode > 3) return -1;
	Joint& j = w->joints[k];
	if (j.kind != 10 && j.kind != 11) return -1;
	// setMotor* / setPiston* (PhysicsEngine.cpp:1683-1912): OFF keeps the
	// parameters; the others set the maximum and their target
	j.mode = mode;
	if (mode >= 1) j.dmax = max;
	if (mode == 2) j.dvel = target;
	if (mode == 3) j.dpos = target;
	const int64_t ends[2] = {j.a, j.b};  // alertWakingAction on both bodies
	orc_wake_bodies(w, 2, ends);
	return 0;
}

void orc_set_hsys(orc_world* w, int n_sys, const int64_t* off, const int64_t* og) {
	w->systems.clear();
	for (Joint& j : w->joints) j.sys = -1;
	for (int k = 0; k < n_sys; k++) {
		HSys hs;
		for (int64_t i = off[k]; i < off[k + 1]; i++) {
			hs.og.push_back((int)og[i]);
			w->joints[og[i]].sys = k;
		}
		hsysLayout(*w, hs);
		w->systems.push_back(hs);
	}
}

int orc_hsys_count(const orc_world* w) { return (int)w->systems.size(); }

int orc_get_joint_rows(const orc_world* w, int k, double* ja, double* jb, double* ita, double* itb) {
	const Joint& j = w->joints[k];
	for (int r = 0; r < j.type; r++)
		for (int c = 0; c < 6; c++) {
			ja[r * 6 + c] = j.ja.v[r][c];
			jb[r * 6 + c] = j.jb.v[r][c];
			ita[c * j.type + r] = j.ita.v[c][r];
			itb[c * j.type + r] = j.itb.v[c][r];
		}
	return j.type;
}

void orc_hsys_prepare(orc_world* w) {
	for (HSys& hs : w->systems) {
		if (!hs.pending) continue;
		for (int ji : hs.og) buildJoint(*w, w->joints[ji], w->cfg.dt / w->cfg.substeps, false);
		hsFactor(*w, hs, w->cfg.cfm);
		hs.pending = false;
	}
}

void orc_get_hsys(const orc_world* w, int k, int64_t* order, int64_t* table, int64_t* vloc, double* buf, int64_t* sizes4) {
	const HSys& hs = w->systems[k];
	sizes4[0] = (int64_t)hs.order.size();
	sizes4[1] = (int64_t)hs.table.size();
	sizes4[2] = hs.degree;
	sizes4[3] = hs.cap;
	if (order) for (size_t i = 0; i < hs.order.size(); i++) order[i] = hs.order[i];
	if (table) for (size_t i = 0; i < hs.table.size(); i++) table[i] = hs.table[i];
	if (vloc) for (size_t i = 0; i < hs.vloc.size(); i++) vloc[i] = hs.vloc[i];
	if (buf) for (int i = 0; i < hs.cap; i++) buf[i] = hs.buf[i];
}

void orc_set_state(orc_world* w, const double* s) {
	for (size_t i = 0; i < w->bodies.size(); i++) {
		Body& b = w->bodies[i];
		const double* r = s + 13 * i;
		b.com = V3(r[0], r[1], r[2]);
		b.q = Q(r[3], r[4], r[5], r[6]);
		b.vel = V3(r[7], r[8], r[9]);
		b.ang_vel = V3(r[10], r[11], r[12]);
		b.pvel = V3();
		b.pang = V3();
	}
}

void orc_get_state(const orc_world* w, double* s) {
	for (size_t i = 0; i < w->bodies.size(); i++) {
		const Body& b = w->bodies[i];
		double* r = s + 13 * i;
		r[0] = b.com.x; r[1] = b.com.y; r[2] = b.com.z;
		r[3] = b.q.r; r[4] = b.q.i; r[5] = b.q.j; r[6] = b.q.k;
		r[7] = b.vel.x; r[8] = b.vel.y; r[9] = b.vel.z;
		r[10] = b.ang_vel.x; r[11] = b.ang_vel.y; r[12] = b.ang_vel.z;
	}
}

// The stored orientation is already normalised by the previous step's
// updateGeometry; the inertia tensor follows it (updateInertiaTensor normalises
// again, which the reference also does every substep).
void orc_refresh(orc_world* w) {
	for (Body& b : w->bodies) {
		Q keep = b.q;
		updateInertia(b);
		b.q = keep;
		worldGeometry(b);
	}
}

void orc_set_inertia(orc_world* w, const double* in) {
	for (size_t i = 0; i < w->bodies.size(); i++) {
		Body& b = w->bodies[i];
		for (int r = 0; r < 3; r++)
			for (int c = 0; c < 3; c++) {
				b.I.v[r][c] = in[18 * i + 3 * r + c];
				b.I_inv.v[r][c] = in[18 * i + 9 + 3 * r + c];
			}
	}
}

void orc_set_cache(orc_world* w, int n, const int64_t* ints, const uint64_t* magic, const double* f) {
	w->cache.clear();
	for (int k = 0; k < n; k++) {
		const int64_t* ii = ints + 6 * k;
		const double* ff = f + 21 * k;
		Contact c{};
		c.b1 = (int)ii[0] - 1;
		c.b2 = (int)ii[1] - 1;
		c.life = (int)ii[2];
		c.live = ii[3] != 0;
		c.f.static_ready = ii[4] != 0;
		c.cid = magic[2 * k];
		c.bid = magic[2 * k + 1];
		c.lp1 = V3(ff[0], ff[1], ff[2]);
		c.lp2 = V3(ff[3], ff[4], ff[5]);
		c.ln1 = V3(ff[6], ff[7], ff[8]);
		c.pen = ff[9];
		c.bounce = ff[10];
		c.mu = ff[11];
		c.c.imp = ff[12];
		c.f.imp[0] = ff[13];
		c.f.imp[1] = ff[14];
		c.f.u = V3(ff[15], ff[16], ff[17]);
		c.f.w = V3(ff[18], ff[19], ff[20]);
		c.hardness = w->cfg.contact_hardness;
		c.cutoff = w->cutoff_vel;
		c.seq = w->seq++;
		w->cache.push_back(c);
	}
}

static void dumpCache(const orc_world* w, const std::vector<Contact>& cache, int64_t* ints, uint64_t* magic, double* f) {
	int pos = 0;
	for (size_t k = 0; k < cache.size(); k++) {
		const Contact& c = cache[k];
		if (k > 0 && (cache[k - 1].b1 != c.b1 || cache[k - 1].b2 != c.b2)) pos = 0;
		int64_t* ii = ints + 6 * k;
		ii[0] = w->bodies[c.b1].id;
		ii[1] = w->bodies[c.b2].id;
		ii[2] = c.life;
		ii[3] = c.live;
		ii[4] = c.f.static_ready;
		ii[5] = pos++;
		magic[2 * k] = c.cid;
		magic[2 * k + 1] = c.bid;
		double* ff = f + 21 * k;
		double v[21] = {c.lp1.x, c.lp1.y, c.lp1.z, c.lp2.x, c.lp2.y, c.lp2.z, c.ln1.x, c.ln1.y, c.ln1.z, c.pen, c.bounce,
		                c.mu, c.c.imp, c.f.imp[0], c.f.imp[1], c.f.u.x, c.f.u.y, c.f.u.z, c.f.w.x, c.f.w.y, c.f.w.z};
		memcpy(ff, v, sizeof v);
	}
}

int orc_cache_size(const orc_world* w) { return (int)w->cache.size(); }
void orc_get_cache(const orc_world* w, int64_t* ints, uint64_t* magic, double* f) { dumpCache(w, w->cache, ints, magic, f); }
int orc_last_mid_cache_size(const orc_world* w) { return (int)w->last_mid.size(); }
void orc_last_mid_cache(const orc_world* w, int64_t* ints, uint64_t* magic, double* f) { dumpCache(w, w->last_mid, ints, magic, f); }

int orc_broadphase(const orc_world* w, int64_t* out, int cap) {
	auto v = broadphase(*w);
	int n = 0;
	for (auto& p : v) {
		if (n < cap) {
			out[2 * n] = w->bodies[p.first].id;
			out[2 * n + 1] = w->bodies[p.second].id;
		}
		n++;
	}
	return n;
}

void orc_get_aabbs(const orc_world* w, double* a) {
	for (size_t i = 0; i < w->bodies.size(); i++) memcpy(a + 6 * i, w->bodies[i].aabb, 6 * sizeof(double));
}

// ---------------------------------------------------------------- raycasts
namespace {
struct RayHit {
	bool hit = false;
	V3 p, n;
	double t = 0;
};

// AABB::rayIntersectsAABB (AABB.h:143-173).
bool rayHitsBox(const double* box, V3 o, V3 d) {
	double t_min = -std::numeric_limits<double>::infinity(), t_max = std::numeric_limits<double>::infinity();
	const double eps = 0.00000000001;
	const double oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
	for (int i = 0; i < 3; i++) {
		if (std::abs(dd[i]) < eps) {
			if (oo[i] < box[i] || oo[i] > box[3 + i]) return false;
		} else {
			double t0 = (box[i] - oo[i]) / dd[i];
			double t1 = (box[3 + i] - oo[i]) / dd[i];
			if (t0 > t1) std::swap(t0, t1);
			t_min = (t0 < t_min) ? t_min : t0;  // std::max<double>(t0, t_min)
			t_max = (t_max < t1) ? t_max : t1;  // std::min<double>(t1, t_max)
			if (t_max < 0 || t_max < t_min) return false;
		}
	}
	return true;
}

// Roots of the ray / quadric quadratic a t^2 + b t + c as the sphere and
// cylinder-side tests take them (ConvexPrimitive.cpp:137-150, :179-192):
// the smaller non-negative root.
bool nearRoot(double a, double b, double c, double* t) {
	double radical = b * b - 4 * a * c;
	if (radical < 0) return false;
	if (radical == 0) {
		*t = -0.5 * b / a;
		return !(*t < 0);
	}
	double t1 = (-b + std::sqrt(radical)) / (2 * a);
	double t2 = (-b - std::sqrt(radical)) / (2 * a);
	if (t1 < 0 && t2 < 0) return false;
	if (t1 < 0) *t = t2;
	else if (t2 < 0) *t = t1;
	else *t = (t2 < t1) ? t2 : t1;  // std::min<double>(t1, t2)
	return true;
}

// testRayAgainstSphere (ConvexPrimitive.cpp:170-202).
RayHit raySphere(V3 o, V3 d, V3 c, double r) {
	RayHit out;
	V3 rel = o - c;
	double a = d.magSqrd(), b = 2 * rel.dot(d), cc = rel.magSqrd() - r * r;
	double t;
	if (!nearRoot(a, b, cc, &t)) return out;
	out.hit = true;
	out.p = o + t * d;
	out.n = (out.p - c).normalize();
	out.t = t;
	return out;
}

// testRayAgainstCylinderSide (ConvexPrimitive.cpp:113-167).
RayHit rayCylinderSide(V3 o, V3 d, V3 c, double radius, double height, V3 ax) {
	RayHit out;
	V3 u = d - ax * ax.dot(d);
	V3 w = o - c - ax * (ax.dot(o) - ax.dot(c));
	double a = u.x * u.x + u.y * u.y + u.z * u.z;
	double b = 2 * (u.x * w.x + u.y * w.y + u.z * w.z);
	double cc = w.x * w.x + w.y * w.y + w.z * w.z - radius * radius;
	double t;
	if (!nearRoot(a, b, cc, &t)) return out;
	V3 hp = o + t * d;
	double x = (hp - c).dot(ax);
	if (std::abs(x) > height / 2.0) return out;
	out.hit = true;
	out.p = hp;
	out.n = ((hp - c) - ax * x).normalize();
	out.t = t;
	return out;
}

// Cylinder::testRayIntersection with checkDiskIntersection (ConvexPrimitive.cpp:300-317).
RayHit rayCylinder(const Body& b, V3 o, V3 d) {
	RayHit out = rayCylinderSide(o, d, b.cyl_c, b.cyl_r, b.cyl_h, b.cyl_ax);
	double sgn = b.cyl_ax.dot(d) < 0 ? 1.0 : -1.0;
	V3 dc = b.cyl_c + (sgn * b.cyl_ax) * (b.cyl_h / 2.0);
	double t = b.cyl_ax.dot(dc - o) / b.cyl_ax.dot(d);
	V3 ip = o + t * d;
	if (!(t < 0 || (ip - dc).magSqrd() > b.cyl_r * b.cyl_r) && (!out.hit || t < out.t)) {
		out.hit = true;
		out.p = ip;
		out.n = b.cyl_ax;
		out.t = t;
	}
	return out;
}

// Polyhedron::testRayIntersection (ConvexPrimitive.cpp:731-750): the first
// surface, in surface order, that the ray enters from outside.
RayHit rayPolyhedron(const Hull& h, V3 o, V3 d) {
	RayHit out;
	for (int f = 0; f < (int)h.faces.size(); f++) {
		const std::vector<int>& fv = h.faces[f];
		V3 sp = h.pts[fv[0]];
		V3 n = faceNormal(h, f);
		const double eps = 0.0001;
		if (d.dot(n) > -eps || (o - sp).dot(n) < 0) continue;
		double t = n.dot(sp - o) / n.dot(d);
		V3 ip = o + t * d;
		bool inside = true;
		int m = (int)fv.size();
		for (int i = 0; i < m && inside; i++) {
			V3 e1 = h.pts[fv[i]], e2 = h.pts[fv[i + 1 == m ? 0 : i + 1]];
			V3 od = (e2 - e1).cross(n);
			if (od.dot(ip - e1) > 0) inside = false;
		}
		if (!inside) continue;
		out.hit = true;
		out.p = ip;
		out.n = n;
		out.t = t;
		return out;
	}
	return out;
}

// StaticMeshGeometry::testRayIntersection (Geometry.cpp:867-895) over the
// candidates of AABBTree::raycastHitCandidates (AABB_Tree.h:249-270): depth
// first, right child popped first, leaves tested on their true boxes.
RayHit rayMesh(const Mesh& M, V3 o, V3 d) {
	RayHit best;
	if (M.root < 0) return best;
	std::vector<int> st{M.root};
	while (!st.empty()) {
		int nd = st.back();
		st.pop_back();
		const BVNode& N = M.nodes[nd];
		if (N.leaf < 0) {
			if (rayHitsBox(N.box, o, d)) {
				st.push_back(N.left);
				st.push_back(N.right);
			}
			continue;
		}
		if (!rayHitsBox(N.tbox, o, d)) continue;
		const MeshTri& tri = M.tris[N.leaf];
		if (std::abs(tri.n.dot(d)) < 0.0000000001) continue;
		double t = -(o - tri.v[1]).dot(tri.n) / d.dot(tri.n);
		if (t < 0 || (best.hit && best.t < t)) continue;
		V3 hp = o + t * d;
		if ((hp - tri.ep1[0]).dot(tri.eout[0]) > 0 || (hp - tri.ep1[1]).dot(tri.eout[1]) > 0 || (hp - tri.ep1[2]).dot(tri.eout[2]) > 0)
			continue;
		best.hit = true;
		best.p = hp;
		best.n = tri.n;
		best.t = t;
	}
	return best;
}

// RigidBody::checkRayIntersection (RigidBody.cpp:185-204) for a body of one
// primitive (its box test is the engine's) or a static mesh.
RayHit rayBody(const Body& b, V3 o, V3 d) {
	switch (b.type) {
	case 0: return rayPolyhedron(b.hull, o, d);
	case 1: return raySphere(o, d, b.sph_c, b.radius);
	case 3: return rayCylinder(b, o, d);
	default: return b.mesh ? rayMesh(*b.mesh, o, d) : RayHit();
	}
}
}  // namespace

// PhysicsEngine::raycastFirstIntersection (PhysicsEngine.cpp:582-609) over the
// bodies [b0, b1) in index order, each candidate a body whose box the ray
// meets; the least distance wins, the first candidate on a tie.
int orc_raycast(const orc_world* w, const double* o3, const double* d3, int b0, int b1, double* out7) {
	V3 o(o3[0], o3[1], o3[2]), d(d3[0], d3[1], d3[2]);
	int best = -1;
	RayHit bh;
	for (int i = b0; i < b1; i++) {
		const Body& b = w->bodies[i];
		if (!rayHitsBox(b.aabb, o, d)) continue;
		RayHit h = rayBody(b, o, d);
		if (h.hit && (best < 0 || h.t < bh.t)) {
			bh = h;
			best = i;
		}
	}
	if (best >= 0) {
		out7[0] = bh.p.x; out7[1] = bh.p.y; out7[2] = bh.p.z;
		out7[3] = bh.n.x; out7[4] = bh.n.y; out7[5] = bh.n.z;
		out7[6] = bh.t;
	}
	return best;
}

void orc_get_world_points(const orc_world* w, int body, double* p) {
	const Hull& h = w->bodies[body].hull;
	for (size_t i = 0; i < h.pts.size(); i++) {
		p[3 * i] = h.pts[i].x; p[3 * i + 1] = h.pts[i].y; p[3 * i + 2] = h.pts[i].z;
	}
}

int orc_narrowphase(const orc_world* w, int id_lo, int id_hi, double* head, double* pts, uint64_t* magic) {
	Manifold m = narrow(*w, id_lo - 1, id_hi - 1);
	if (m.pts.empty()) return 0;
	head[0] = m.n.x; head[1] = m.n.y; head[2] = m.n.z; head[3] = m.max_pen;
	for (size_t k = 0; k < m.pts.size(); k++) {
		pts[4 * k] = m.pts[k].pos.x; pts[4 * k + 1] = m.pts[k].pos.y; pts[4 * k + 2] = m.pts[k].pos.z; pts[4 * k + 3] = m.pts[k].pen;
		magic[2 * k] = m.pts[k].cid;
		magic[2 * k + 1] = m.pts[k].bid;
	}
	return (int)m.pts.size();
}

void orc_step(orc_world* w, const int64_t* node_order) { step(*w, node_order); }

int orc_step_scheduled(orc_world* w, int n_groups, const int64_t* rows, const int64_t* offs) { return step(*w, nullptr, rows, offs, n_groups); }

void orc_set_sleeping(orc_world* w, int enabled, const double* params4, int tick_threshold) {
	w->sleeping = enabled;
	if (params4) {
		w->sleep_delay = params4[0];
		w->vel_sleep_coeff = params4[1];
		w->ang_vel_eps = params4[2];
		w->accel_sleep_coeff = params4[3];
	}
	if (tick_threshold > 0) w->non_sleepy_tick_threshold = tick_threshold;
}

void orc_get_sleep_state(const orc_world* w, double* out4) {
	for (size_t i = 0; i < w->bodies.size(); i++) {
		const Body& b = w->bodies[i];
		out4[4 * i] = b.asleep ? 1.0 : 0.0;
		out4[4 * i + 1] = b.sleep_counter;
		out4[4 * i + 2] = (double)b.sleep_ticks;
		out4[4 * i + 3] = (double)b.history.size();
	}
}

void orc_wake_bodies(orc_world* w, int n, const int64_t* idx) {
	for (int k = 0; k < n; k++) {
		Body& b = w->bodies[idx[k]];
		// RigidBody::alertWakingAction (RigidBody.cpp:340-345)
		b.wake_req = true;
		b.sleep_counter = 0;
		b.sleep_ticks = 0;
		b.history.clear();
	}
}

void orc_set_sleep_disabled(orc_world* w, int n, const int64_t* idx, int disabled) {
	for (int k = 0; k < n; k++) w->bodies[idx[k]].sleep_disabled = disabled != 0;
}

int orc_last_islands(const orc_world* w, int64_t* rows, int64_t* offs, int cap) {
	int n = (int)w->last_rows.size() / 2;
	if (n <= cap) {
		memcpy(rows, w->last_rows.data(), w->last_rows.size() * sizeof(int64_t));
		memcpy(offs, w->last_offs.data(), w->last_offs.size() * sizeof(int64_t));
	}
	return (int)w->last_offs.size();
}

}  // extern "C"
