// Host-side shape construction.  Follows the reference's Polyhedron
// constructor so that shapes built here carry the same surfaces, winding,
// normal-evaluation vertices, edges and Gauss map (arcs between faces that
// share an edge, antipodal-face redundancy flags, cached self-extrema).
#include "shapes.h"

#include <cmath>
#include <limits>
#include <unordered_map>
#include <utility>
#include <vector>

#include "cp_math.h"

namespace cpx {

int HostShapes::add_hull(const HullData& d) {
	int id = (int)type.size();
	type.push_back(0);
	int p0 = pt_off.back();
	int np = (int)d.pts.size() / 3;
	ref_pts.insert(ref_pts.end(), d.pts.begin(), d.pts.end());
	pt_off.push_back(p0 + np);
	std::vector<std::vector<int>> vf(np), ve(np);
	for (size_t f = 0; f < d.faces.size(); f++) {
		for (int v : d.faces[f]) {
			fv.push_back(v);
			vf[v].push_back((int)f);
		}
		fv_off.push_back((int)fv.size());
		ncalc.push_back(d.ncalc[f]);
		ndir.push_back(d.ndir[f]);
		sid.push_back(d.sid[f]);
		for (int c = 0; c < 3; c++) gnorm.push_back(d.gnorm[3 * f + c]);
		red.push_back(d.red[f]);
		minpid.push_back(d.minpid[f]);
		maxpid.push_back(d.maxpid[f]);
		refidx.push_back(d.refidx[f]);
		gmin.push_back(d.gmin[f]);
		gmax.push_back(d.gmax[f]);
		grefval.push_back(d.grefval[f]);
	}
	face_off.push_back(face_off.back() + (int)d.faces.size());
	for (auto& a : d.arcs) { arcs.push_back(a.first); arcs.push_back(a.second); }
	arc_off.push_back(arc_off.back() + (int)d.arcs.size());
	for (size_t e = 0; e < d.edges.size(); e++) {
		edges.push_back(d.edges[e].first);
		edges.push_back(d.edges[e].second);
		ve[d.edges[e].first].push_back((int)e);
		ve[d.edges[e].second].push_back((int)e);
	}
	edge_off.push_back(edge_off.back() + (int)d.edges.size());
	for (int v = 0; v < np; v++) {
		vaf.insert(vaf.end(), vf[v].begin(), vf[v].end());
		vaf_off.push_back((int)vaf.size());
		vae.insert(vae.end(), ve[v].begin(), ve[v].end());
		vae_off.push_back((int)vae.size());
	}
	radius.push_back(0);
	height.push_back(0);
	return id;
}

int HostShapes::add_sphere(const double* center, double r) {
	int id = (int)type.size();
	type.push_back(1);
	// the centre in the body frame is the shape's single reference point
	ref_pts.insert(ref_pts.end(), center, center + 3);
	pt_off.push_back(pt_off.back() + 1);
	vaf_off.push_back((int)vaf.size());
	vae_off.push_back((int)vae.size());
	face_off.push_back(face_off.back());
	arc_off.push_back(arc_off.back());
	edge_off.push_back(edge_off.back());
	radius.push_back(r);
	height.push_back(0);
	return id;
}

int HostShapes::add_cylinder(const double* params8, const double* gauss42, const double* top24, const double* bot24) {
	int id = (int)type.size();
	type.push_back(2);
	ref_pts.insert(ref_pts.end(), params8, params8 + 3);
	ref_pts.insert(ref_pts.end(), top24, top24 + 24);
	ref_pts.insert(ref_pts.end(), bot24, bot24 + 24);
	pt_off.push_back(pt_off.back() + 17);
	for (int v = 0; v < 17; v++) {
		vaf_off.push_back((int)vaf.size());
		vae_off.push_back((int)vae.size());
	}
	// 15 normal slots (axis, Gauss vertices); the hull face fields stay empty
	for (int f = 0; f < 15; f++) {
		const double* g = f == 0 ? params8 + 3 : gauss42 + 3 * (f - 1);
		gnorm.insert(gnorm.end(), {g[0], g[1], g[2]});
		fv_off.push_back((int)fv.size());
		ncalc.push_back(0);
		ndir.push_back(0);
		sid.push_back(0);
		red.push_back(1);
		minpid.push_back(0);
		maxpid.push_back(0);
		refidx.push_back(0);
		gmin.push_back(0);
		gmax.push_back(0);
		grefval.push_back(0);
	}
	face_off.push_back(face_off.back() + 15);
	arc_off.push_back(arc_off.back());
	edge_off.push_back(edge_off.back());
	radius.push_back(params8[6]);
	height.push_back(params8[7]);
	return id;
}

void build_cylinder(const double* center, double radius, double height, double* params8, double* gauss42, double* top24, double* bot24) {
	const double PI = 3.1415926535;  // the reference's PI macro
	params8[0] = center[0]; params8[1] = center[1]; params8[2] = center[2];
	params8[3] = 0; params8[4] = 1; params8[5] = 0;
	params8[6] = radius;
	params8[7] = height;
	double g[14][3] = {{0, 1, 0}, {0, -1, 0}};
	double dt = 2 * PI / 12;
	for (int i = 0; i < 12; i++) {
		double th = i * dt;
		g[i + 2][0] = std::sin(th); g[i + 2][1] = 0; g[i + 2][2] = std::cos(th);
	}
	for (int i = 0; i < 14; i++)
		for (int c = 0; c < 3; c++) gauss42[3 * i + c] = g[i][c];
	double df = 2 * PI / 8;
	double er = radius / std::cos(PI / 8);
	for (int i = 0; i < 8; i++) {
		double th = i * df;
		double x = er * std::sin(th), z = er * std::cos(th);
		top24[3 * i] = center[0] + x; top24[3 * i + 1] = center[1] + height / 2.0; top24[3 * i + 2] = center[2] + z;
		bot24[3 * i] = center[0] + x; bot24[3 * i + 1] = center[1] + -height / 2.0; bot24[3 * i + 2] = center[2] + z;
	}
}

namespace {

struct Surf {
	std::vector<int> idx;
	int ncalc = 2;
	int ndir = 1;
};

V3 P(const std::vector<double>& p, int i) { return V3{p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }

V3 surf_normal(const std::vector<double>& pts, const Surf& s) {  // ConvexPrimitive.cpp:815-821
	V3 p1 = P(pts, s.idx[0]), p2 = P(pts, s.idx[1]), p3 = P(pts, s.idx[s.ncalc]);
	return normalize(cross(p2 - p1, p3 - p2)) * (double)s.ndir;
}

void find_normal_calc(const std::vector<double>& pts, Surf& s, V3 normalish) {  // ConvexPrimitive.cpp:791-808
	V3 p0 = P(pts, s.idx[0]), p1 = P(pts, s.idx[1]);
	s.ncalc = 2;
	V3 v1 = normalize(p1 - p0);
	V3 v2 = normalize(P(pts, s.idx[s.ncalc]) - p0);
	while (std::abs(dot(v1, v2)) > 0.99 && s.ncalc + 1 < (int)s.idx.size()) v2 = normalize(P(pts, s.idx[++s.ncalc]) - p0);
	s.ndir = 1;
	if (dot(normalish, surf_normal(pts, s)) < 0) s.ndir = -1;
}

void set_winding_ccw(const std::vector<double>& pts, Surf& s, V3 n) {  // ConvexPrimitive.cpp:834-858
	V3 u, w;
	perp_basis(n, &u, &w);
	double area = 0;
	int m = (int)s.idx.size();
	for (int i = 0; i < m; i++) {
		int i2 = (i + 1 == m) ? 0 : i + 1;
		V3 a = P(pts, s.idx[i]), b = P(pts, s.idx[i2]);
		double au = dot(a, u), aw = dot(a, w), bu = dot(b, u), bw = dot(b, w);
		area += (bw - aw) * (bu + au) / 2.0;
	}
	if (area < 0) {
		for (int i = 0; i < m / 2; i++) std::swap(s.idx[i], s.idx[m - 1 - i]);
		find_normal_calc(pts, s, n);
	}
}

}  // namespace

HullData build_polyhedron(const std::vector<double>& pts, const std::vector<std::vector<int>>& faces) {
	HullData d;
	d.pts = pts;
	int np = (int)pts.size() / 3;
	V3 interior{0, 0, 0};
	for (int i = 0; i < np; i++) interior += P(pts, i);
	interior = V3{interior.x / (double)np, interior.y / (double)np, interior.z / (double)np};
	std::vector<Surf> surfs;
	for (size_t f = 0; f < faces.size(); f++) {
		Surf s;
		s.idx = faces[f];
		find_normal_calc(pts, s, P(pts, s.idx[0]) - interior);
		set_winding_ccw(pts, s, surf_normal(pts, s));
		surfs.push_back(s);
		d.faces.push_back(s.idx);
		d.sid.push_back(np + (int)f);
		d.ncalc.push_back(s.ncalc);
		d.ndir.push_back(s.ndir);
	}
	// edges in first-seen order over the (winding-fixed) surfaces
	for (const Surf& s : surfs) {
		int m = (int)s.idx.size();
		for (int i = 0; i < m; i++) {
			int a = s.idx[i], b = s.idx[(i + 1) % m];
			bool seen = false;
			for (auto& e : d.edges)
				if ((e.first == a && e.second == b) || (e.first == b && e.second == a)) { seen = true; break; }
			if (!seen) d.edges.push_back({a, b});
		}
	}
	// Gauss map
	std::vector<V3> gv;
	for (size_t i = 0; i < surfs.size(); i++) {
		V3 n = surf_normal(pts, surfs[i]);
		bool redundant = false;
		for (V3 g : gv)
			if (dot(n, g) < -1 + 0.0000000001) { redundant = true; break; }
		gv.push_back(n);
		d.gnorm.insert(d.gnorm.end(), {n.x, n.y, n.z});
		d.red.push_back(redundant);
		int mn = -1, mx = -1;
		double vmin = std::numeric_limits<double>::infinity(), vmax = -std::numeric_limits<double>::infinity();
		for (int k = 0; k < np; k++) {
			double v = dot(P(pts, k), n);
			if (v < vmin) { mn = k; vmin = v; }
			if (v > vmax) { mx = k; vmax = v; }
		}
		int ref = surfs[i].idx[0];
		d.minpid.push_back(mn);
		d.maxpid.push_back(mx);
		d.refidx.push_back(ref);
		d.gmin.push_back(vmin);
		d.gmax.push_back(vmax);
		d.grefval.push_back(dot(n, P(pts, ref)));
		for (size_t j = i + 1; j < surfs.size(); j++) {
			const Surf& s1 = surfs[i];
			const Surf& s2 = surfs[j];
			int n1 = (int)s1.idx.size(), n2 = (int)s2.idx.size();
			for (int k = 0; k < n1; k++) {
				for (int w = 0; w < n2; w++) {
					int a1 = s1.idx[k], a2 = s1.idx[(k + 1) % n1], b1 = s2.idx[w], b2 = s2.idx[(w + 1) % n2];
					if ((a1 == b1 && a2 == b2) || (a1 == b2 && a2 == b1)) {
						d.arcs.push_back({(int)i, (int)j});
						break;
					}
				}
			}
		}
	}
	return d;
}

namespace {
// getAxisVal (RigidBody.cpp:368-402): the projection plane's axes A, B and
// its normal axis C for a surface projected along X, Y or Z.
enum ProjAxis { PX, PY, PZ };
double axis_val(V3 v, int ax, ProjAxis p) {
	static const int map[3][3] = {{1, 2, 0}, {2, 0, 1}, {0, 1, 2}};
	int c = map[p][ax];
	return c == 0 ? v.x : c == 1 ? v.y : v.z;
}
struct Intrg {
	double v = 0, v2 = 0, v2t = 0, v3 = 0;
};
}  // namespace

// calculateMassProperties for one POLYHEDRON primitive of a convex union
// (RigidBody.cpp:427-532, then :633-637): projection integrals per surface
// after Mirtich, in the reference's operation order, over the surfaces as the
// Polyhedron constructor leaves them (winding fixed, normal from
// Surface::normal).  The tensor is returned about the centre of mass
// (recenterTensor, RigidBody.cpp:406-425).
void polyhedron_mass_properties(const std::vector<double>& pts, const std::vector<std::vector<int>>& faces, double density,
                                double* mass, double* tensor9, double* com3) {
	HullData d = build_polyhedron(pts, faces);
	double vol = 0, vol_x = 0, vol_y = 0, vol_z = 0, vol_xy = 0, vol_yz = 0, vol_zx = 0, vol_x2 = 0, vol_y2 = 0, vol_z2 = 0;
	for (size_t f = 0; f < d.faces.size(); f++) {
		const std::vector<int>& s = d.faces[f];
		V3 n{d.gnorm[3 * f], d.gnorm[3 * f + 1], d.gnorm[3 * f + 2]};  // Surface::normal()
		ProjAxis pa;
		if (std::abs(n.z) > 0.01) pa = PZ;
		else if (std::abs(n.x) >= std::abs(n.y) && std::abs(n.x) >= std::abs(n.z)) pa = PX;
		else if (std::abs(n.y) >= std::abs(n.x) && std::abs(n.y) >= std::abs(n.z)) pa = PY;
		else pa = PZ;
		double g1 = 0, ga = 0, gb = 0, gab = 0, ga2 = 0, gb2 = 0, ga2b = 0, gab2 = 0, ga3 = 0, gb3 = 0;
		int m = (int)s.size();
		for (int i = 0; i < m; i++) {
			V3 v = P(pts, s[i]), v2 = P(pts, s[(i + 1) % m]);
			double va = axis_val(v, 0, pa), vb = axis_val(v, 1, pa);
			double dA = axis_val(v2, 0, pa) - va, dB = axis_val(v2, 1, pa) - vb;
			g1 += dB * (va + dA / 2.0);
			ga += dB * (dA * dA / 3.0 + va * dA + va * va) / 2.0;
			gb -= dA * (dB * dB / 3.0 + vb * dB + vb * vb) / 2.0;
			gab -= dA * (va * vb * vb + dA * vb * vb / 2.0 + va * vb * dB + (va * dB * dB + 2 * vb * dA * dB) / 3.0 + dA * dB * dB / 4.0) / 2.0;
			ga2 += dB * (va * va * va + 3 * va * va * dA / 2.0 + va * dA * dA + dA * dA * dA / 4.0) / 3.0;
			gb2 -= dA * (vb * vb * vb + 3 * vb * vb * dB / 2.0 + vb * dB * dB + dB * dB * dB / 4.0) / 3.0;
			double mixed = va * va * vb * vb + va * va * vb * dB + va * vb * vb * dA + (va * va * dB * dB + vb * vb * dA * dA + 4 * va * vb * dA * dB) / 3.0 +
			               (va * dA * dB * dB + vb * dB * dA * dA) / 2.0 + dA * dA * dB * dB / 5.0;
			ga2b -= dA * mixed / 2.0;
			gab2 += dB * mixed / 2.0;
			ga3 += dB * (va * va * va * va + 2 * va * va * va * dA + 2 * va * va * dA * dA + va * dA * dA * dA + dA * dA * dA * dA / 5.0) / 4.0;
			gb3 -= dA * (vb * vb * vb * vb + 2 * vb * vb * vb * dB + 2 * vb * vb * dB * dB + vb * dB * dB * dB + dB * dB * dB * dB / 5.0) / 4.0;
		}
		double na = axis_val(n, 0, pa), nb = axis_val(n, 1, pa), nc = axis_val(n, 2, pa);
		double nc_i = 1.0 / nc;
		V3 sp = P(pts, s[0]);
		double k = nc * axis_val(sp, 2, pa) + nb * axis_val(sp, 1, pa) + na * axis_val(sp, 0, pa);
		Intrg a, b, c;
		a.v = nc_i * ga;
		b.v = nc_i * gb;
		c.v = nc_i * nc_i * (k * g1 - na * ga - nb * gb);
		a.v2 = nc_i * ga2;
		b.v2 = nc_i * gb2;
		c.v2 = nc_i * nc_i * nc_i * (k * k * g1 + na * na * ga2 + nb * nb * gb2 + 2 * (na * nb * gab - k * na * ga - k * nb * gb));
		a.v2t = nc_i * ga2b;
		b.v2t = nc_i * nc_i * (k * gb2 - na * gab2 - nb * gb3);
		c.v2t = nc_i * nc_i * nc_i * (k * k * ga + na * na * ga3 + nb * nb * gab2 + 2 * (na * nb * ga2b - k * na * ga2 - k * nb * gab));
		a.v3 = nc_i * ga3;
		b.v3 = nc_i * gb3;
		c.v3 = nc_i * nc_i * nc_i * nc_i *
		       (k * k * k * g1 - na * na * na * ga3 - nb * nb * nb * gb3 - 3 * k * k * na * ga + 3 * k * na * na * ga2 - 3 * na * na * nb * ga2b -
		        3 * na * nb * nb * gab2 + 3 * k * nb * nb * gb2 - 3 * k * k * nb * gb + 6 * k * na * nb * gab);
		Intrg x, y, z;
		if (pa == PX) { x = c; y = a; z = b; }
		else if (pa == PY) { x = b; y = c; z = a; }
		else { x = a; y = b; z = c; }
		vol += n.x * x.v;
		vol_x += n.x * x.v2 / 2.0;
		vol_y += n.y * y.v2 / 2.0;
		vol_z += n.z * z.v2 / 2.0;
		vol_xy += n.x * x.v2t / 2.0;
		vol_yz += n.y * y.v2t / 2.0;
		vol_zx += n.z * z.v2t / 2.0;
		vol_x2 += n.x * x.v3 / 3.0;
		vol_y2 += n.y * y.v3 / 3.0;
		vol_z2 += n.z * z.v3 / 3.0;
	}
	double ms = 0;
	ms += density * vol;
	V3 com{0, 0, 0};
	com += V3{vol_x, vol_y, vol_z} * density;
	double t[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
	t[0][0] += (vol_y2 + vol_z2) * density;
	t[0][1] -= vol_xy * density;
	t[0][2] -= vol_zx * density;
	t[1][1] += (vol_x2 + vol_z2) * density;
	t[1][2] -= vol_yz * density;
	t[2][2] += (vol_x2 + vol_y2) * density;
	com = V3{com.x / ms, com.y / ms, com.z / ms};
	// recenterTensor(mass, tensor, com, 0, com)
	V3 dd = com - V3{0, 0, 0}, cc = com - V3{0, 0, 0};
	t[0][0] += ms * (dd.y * dd.y + dd.z * dd.z - 2 * dd.y * cc.y - 2 * dd.z * cc.z);
	t[0][1] += ms * (dd.x * cc.y + dd.y * cc.x - dd.x * dd.y);
	t[0][2] += ms * (dd.x * cc.z + dd.z * cc.x - dd.x * dd.z);
	t[1][1] += ms * (dd.x * dd.x + dd.z * dd.z - 2 * dd.x * cc.x - 2 * dd.z * cc.z);
	t[1][2] += ms * (dd.y * cc.z + dd.z * cc.y - dd.y * dd.z);
	t[2][2] += ms * (dd.x * dd.x + dd.y * dd.y - 2 * dd.x * cc.x - 2 * dd.y * cc.y);
	t[1][0] = t[0][1];
	t[2][0] = t[0][2];
	t[2][1] = t[1][2];
	*mass = ms;
	for (int i = 0; i < 3; i++)
		for (int j = 0; j < 3; j++) tensor9[3 * i + j] = t[i][j];
	com3[0] = com.x;
	com3[1] = com.y;
	com3[2] = com.z;
}

// Mat3::inverse (Mat3.h:57-68): cofactors divided by the determinant.
void mat3_inverse(const double* m9, double* out9) {
	double a = m9[0], b = m9[1], c = m9[2], d = m9[3], e = m9[4], f = m9[5], g = m9[6], h = m9[7], i = m9[8];
	double det = a * (e * i - f * h) - b * (d * i - f * g) + c * (d * h - e * g);
	double o[9] = {(e * i - f * h), (c * h - b * i), (b * f - c * e), (f * g - d * i), (a * i - c * g), (c * d - a * f),
	               (d * h - e * g), (b * g - a * h), (a * e - b * d)};
	for (int k = 0; k < 9; k++) out9[k] = o[k] / det;
}

// A dynamic body built from Polyhedron(points, faces) (RigidBody ctor,
// RigidBody.cpp:13-38): mass properties of the points as given, the Gauss map
// (with its cached self-extrema) of the points as given, and the body-frame
// points recomputeFromReference(identity, -com) leaves: iden * p + (-com).
HullData polyhedron_body(const std::vector<double>& pts, const std::vector<std::vector<int>>& faces, double density, double* mp19) {
	double com[3];
	polyhedron_mass_properties(pts, faces, density, mp19, mp19 + 1, com);
	mat3_inverse(mp19 + 1, mp19 + 10);
	HullData d = build_polyhedron(pts, faces);
	for (size_t i = 0; i + 2 < d.pts.size(); i += 3) {
		double x = pts[i], y = pts[i + 1], z = pts[i + 2];
		double r[3] = {1 * x + 0 * y + 0 * z, 0 * x + 1 * y + 0 * z, 0 * x + 0 * y + 1 * z};
		for (int c = 0; c < 3; c++) d.pts[i + c] = -com[c] + r[c];
	}
	for (size_t f = 0; f + 2 < d.gnorm.size(); f += 3) {
		double x = d.gnorm[f], y = d.gnorm[f + 1], z = d.gnorm[f + 2];
		d.gnorm[f] = 1 * x + 0 * y + 0 * z;
		d.gnorm[f + 1] = 0 * x + 1 * y + 0 * z;
		d.gnorm[f + 2] = 0 * x + 0 * y + 1 * z;
	}
	return d;
}

}  // namespace cpx

// ---------------------------------------------------------------- static meshes
namespace cpx {

int HostShapes::add_mesh(HostMesh m) {
	int id = (int)type.size();
	type.push_back(3);
	pt_off.push_back(pt_off.back());
	face_off.push_back(face_off.back());
	arc_off.push_back(arc_off.back());
	edge_off.push_back(edge_off.back());
	radius.push_back(0);
	height.push_back(0);
	mesh_of.resize(type.size(), -1);
	mesh_of[id] = (int)meshes.size();
	meshes.push_back(std::move(m));
	return id;
}

// StaticMeshGeometry(const MeshInput&) (Geometry.cpp:678-820).  Edges are
// keyed (max index << 32) + min index; the second triangle met on a key
// becomes the neighbour of the first (and the key is retired).  Each triangle
// takes three fresh edge ids from T + P on (the ids are assigned on a copy of
// the graph node, so shared edges get one id per side).  An edge's Gauss
// region contribution is its outward direction on a border, the triangle's
// own normal once for concave edges, else the neighbour's normal.
HostMesh build_mesh(int n_pts, const double* pts, int n_tris, const int64_t* tri_idx, const double* tri_mat3) {
	auto P = [&](int i) { return V3{pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]}; };
	struct Side {
		int a, b, opp;
		uint64_t key;
	};
	struct Node {
		int nb[3];
		Side e[3];
		V3 n;
	};
	std::vector<Node> g(n_tris);
	for (int t = 0; t < n_tris; t++) {
		int i[3] = {(int)tri_idx[3 * t], (int)tri_idx[3 * t + 1], (int)tri_idx[3 * t + 2]};
		g[t].n = normalize(cross(P(i[1]) - P(i[0]), P(i[2]) - P(i[0])));
		for (int k = 0; k < 3; k++) {
			int a = i[k], b = i[(k + 1) % 3], o = i[(k + 2) % 3];
			uint64_t hi = (uint64_t)(a > b ? a : b), lo = (uint64_t)(a > b ? b : a);
			g[t].e[k] = Side{a, b, o, ((hi << 32) & 0xFFFFFFFF00000000ull) + lo};
			g[t].nb[k] = -1;
		}
	}
	std::unordered_map<uint64_t, int> open;
	for (int t = 0; t < n_tris; t++)
		for (int k = 0; k < 3; k++) {
			auto it = open.find(g[t].e[k].key);
			if (it == open.end()) {
				open[g[t].e[k].key] = t;
				continue;
			}
			int o = it->second;
			g[t].nb[k] = o;
			for (int q = 0; q < 3; q++)
				if (g[o].e[q].key == g[t].e[k].key) g[o].nb[q] = t;
			open.erase(it);
		}
	HostMesh m;
	m.tris.resize(n_tris);
	int next_edge = n_tris + n_pts;
	for (int t = 0; t < n_tris; t++) {
		const Node& nd = g[t];
		MeshTriangle& r = m.tris[t];
		V3 v[3] = {P((int)tri_idx[3 * t]), P((int)tri_idx[3 * t + 1]), P((int)tri_idx[3 * t + 2])};
		r.n[0] = nd.n.x; r.n[1] = nd.n.y; r.n[2] = nd.n.z;
		for (int k = 0; k < 3; k++) {
			r.v[3 * k] = v[k].x; r.v[3 * k + 1] = v[k].y; r.v[3 * k + 2] = v[k].z;
			r.mat[k] = tri_mat3[3 * t + k];
		}
		r.ngauss = 0;
		r.concave = 0;
		auto push_gauss = [&](V3 d) {
			r.gauss[3 * r.ngauss] = d.x; r.gauss[3 * r.ngauss + 1] = d.y; r.gauss[3 * r.ngauss + 2] = d.z;
			r.ngauss++;
		};
		for (int k = 0; k < 3; k++) {
			V3 out = normalize(cross(v[(k + 1) % 3] - v[k], nd.n));
			r.eout[3 * k] = out.x; r.eout[3 * k + 1] = out.y; r.eout[3 * k + 2] = out.z;
			r.eid[k] = next_edge++;
			if (nd.nb[k] < 0) {
				push_gauss(out);
				continue;
			}
			const Node& o = g[nd.nb[k]];
			int opp = nd.e[k].opp;
			for (int q = 0; q < 3; q++)
				if (o.e[q].a == nd.e[k].b && o.e[q].b == nd.e[k].a) opp = o.e[q].opp;
			const double EPS = 0.00001;
			if (dot(normalize(P(opp) - P(nd.e[k].opp)), nd.n) >= -EPS) {
				r.concave++;
				if (r.concave <= 1) push_gauss(nd.n);
			} else {
				push_gauss(o.n);
			}
		}
	}
	return m;
}

namespace {
// AABB::combine / surfaceArea (AABB.h:16-26, :112-118, as written).
void box_combine(const double* a, const double* b, double* o) {
	for (int i = 0; i < 3; i++) {
		o[i] = (b[i] < a[i]) ? b[i] : a[i];
		o[3 + i] = (a[3 + i] < b[3 + i]) ? b[3 + i] : a[3 + i];
	}
}
double box_cost(const double* a) {
	double dx = a[3] - a[0], dy = a[4] - a[1], dz = a[5] - a[2];
	return 2 * (dx * dy) + (dy * dz) + (dz * dx);
}
struct TreeBuild {
	std::vector<double> box, cost;
	std::vector<int> parent, left, right, leaf;
	int root = -1;
	int add_node() {
		int n = (int)parent.size();
		box.resize(6 * (n + 1));
		cost.push_back(0);
		parent.push_back(-1);
		left.push_back(-1);
		right.push_back(-1);
		leaf.push_back(-1);
		return n;
	}
	void set_box(int n, const double* b) {
		for (int i = 0; i < 6; i++) box[6 * n + i] = b[i];
		cost[n] = box_cost(b);
	}
	void swap_nodes(int n1, int n2) {  // Node::swpNodes (AABB_Tree.h:306-319)
		int p1 = parent[n1], p2 = parent[n2];
		int* s1 = left[p1] == n1 ? &left[p1] : &right[p1];
		int* s2 = left[p2] == n2 ? &left[p2] : &right[p2];
		*s1 = n2;
		*s2 = n1;
		parent[n1] = p2;
		parent[n2] = p1;
	}
	// AABBTree::add for a static object (AABB_Tree.h:39-103): branch and bound
	// over insertion siblings, a new parent, box refresh up the tree, then
	// rebalanceAncestors (:377-424).
	void add(int tri, const double* tb) {
		int nl = add_node();
		leaf[nl] = tri;
		double mb[6];
		for (int i = 0; i < 3; i++) {
			mb[i] = tb[i] - 0.0;
			mb[3 + i] = tb[3 + i] + 0.0;
		}
		set_box(nl, mb);
		if (root < 0) {
			root = nl;
			return;
		}
		int best = -1;
		double best_cost = std::numeric_limits<double>::infinity();
		std::vector<std::pair<int, double>> cand{{root, 0.0}};
		while (!cand.empty()) {
			std::pair<int, double> c = cand.back();
			cand.pop_back();
			if (c.second >= best_cost) continue;
			double cb[6];
			box_combine(&box[6 * nl], &box[6 * c.first], cb);
			double direct = box_cost(cb);
			if (c.second + direct < best_cost) {
				best = c.first;
				best_cost = c.second + direct;
			}
			if (leaf[c.first] < 0) {
				double inherited = c.second + direct - cost[c.first];
				cand.push_back({left[c.first], inherited});
				cand.push_back({right[c.first], inherited});
			}
		}
		int np = add_node();
		if (best == root) {
			root = np;
		} else {
			int op = parent[best];
			if (left[op] == best) left[op] = np;
			else right[op] = np;
			parent[np] = op;
		}
		left[np] = best;
		parent[best] = np;
		right[np] = nl;
		parent[nl] = np;
		for (int c = parent[nl]; c >= 0; c = parent[c]) {
			double cb[6];
			box_combine(&box[6 * left[c]], &box[6 * right[c]], cb);
			set_box(c, cb);
		}
		for (int c = parent[nl]; c >= 0 && parent[c] >= 0; c = parent[c]) {
			int par = parent[c];
			int sib = left[par] == c ? right[par] : left[par];
			double sl[6], sr[6], cl[6], cr[6], clc, crc;
			box_combine(&box[6 * right[c]], &box[6 * sib], sl);
			box_combine(&box[6 * left[c]], &box[6 * sib], sr);
			double slc = box_cost(sl), src = box_cost(sr);
			if (leaf[sib] >= 0) {
				clc = std::numeric_limits<double>::infinity();
				crc = clc;
			} else {
				box_combine(&box[6 * right[sib]], &box[6 * c], cl);
				box_combine(&box[6 * left[sib]], &box[6 * c], cr);
				clc = box_cost(cl);
				crc = box_cost(cr);
			}
			double dcl = clc - cost[sib], dcr = crc - cost[sib], dsl = slc - cost[c], dsr = src - cost[c];
			double m1 = (dsr < dsl) ? dsr : dsl;
			double m2 = (m1 < dcr) ? m1 : dcr;
			double mn = (m2 < dcl) ? m2 : dcl;  // std::min<double>(cl, min(cr, min(sl, sr)))
			if (mn >= 0) continue;
			if (mn == dcl) {
				swap_nodes(c, left[sib]);
				for (int i = 0; i < 6; i++) box[6 * sib + i] = cl[i];
				cost[sib] = clc;
			} else if (mn == dcr) {
				swap_nodes(c, right[sib]);
				for (int i = 0; i < 6; i++) box[6 * sib + i] = cr[i];
				cost[sib] = crc;
			} else if (mn == dsl) {
				swap_nodes(sib, left[c]);
				for (int i = 0; i < 6; i++) box[6 * c + i] = sl[i];
				cost[c] = slc;
			} else {
				swap_nodes(sib, right[c]);
				for (int i = 0; i < 6; i++) box[6 * c + i] = sr[i];
				cost[c] = src;
			}
		}
	}
};
}  // namespace

// recomputeFromReference with the body's pose (getTransformed, Geometry.cpp:
// 827-847, centre of rotation 0), the AABB tree over the world triangles and
// genAABB.  Leaves keep the triangle's exact box; the stack depth of a query
// is recorded for the device traversal.
WorldMesh place_mesh(const HostMesh& m, const double* com3, const double* q4) {
	Q q = q_normalize(Q{q4[0], q4[1], q4[2], q4[3]});  // RigidBody::updateGeometry (RigidBody.cpp:283)
	M3 R = q_rot(q);
	V3 t{com3[0], com3[1], com3[2]}, c{0, 0, 0};
	int T = (int)m.tris.size();
	WorldMesh w;
	w.tri.assign((size_t)T * MT_REC, 0.0);
	w.tri_i.assign((size_t)T * MT_IREC, 0);
	auto ld = [](const double* a, int k) { return V3{a[3 * k], a[3 * k + 1], a[3 * k + 2]}; };
	auto st = [](double* a, int k, V3 v) { a[3 * k] = v.x; a[3 * k + 1] = v.y; a[3 * k + 2] = v.z; };
	TreeBuild tb;
	for (int i = 0; i < T; i++) {
		const MeshTriangle& r = m.tris[i];
		double* o = &w.tri[(size_t)i * MT_REC];
		st(o, 0, R * ld(r.n, 0));
		const double inf = std::numeric_limits<double>::infinity();
		double mn[3] = {inf, inf, inf}, mx[3] = {-inf, -inf, -inf};
		for (int k = 0; k < 3; k++) {
			V3 v = t + R * (ld(r.v, k) - c) + c;
			st(o + 3, k, v);
			st(o + 12, k, R * ld(r.eout, k));
			double p[3] = {v.x, v.y, v.z};
			for (int d = 0; d < 3; d++) {
				mn[d] = (p[d] < mn[d]) ? p[d] : mn[d];
				mx[d] = (mx[d] < p[d]) ? p[d] : mx[d];
			}
		}
		for (int k = 0; k < r.ngauss; k++) st(o + 21, k, R * ld(r.gauss, k));
		for (int d = 0; d < 3; d++) {
			o[30 + d] = mn[d];
			o[33 + d] = mx[d];
			o[36 + d] = r.mat[d];
		}
		int* oi = &w.tri_i[(size_t)i * MT_IREC];
		for (int k = 0; k < 3; k++) oi[k] = r.eid[k];
		oi[3] = r.ngauss;
		oi[4] = r.concave;
		oi[5] = i;
		tb.add(i, o + 30);
		if (i == 0) for (int d = 0; d < 6; d++) w.aabb[d] = o[30 + d];
		else {
			double cb[6];
			box_combine(w.aabb, o + 30, cb);
			for (int d = 0; d < 6; d++) w.aabb[d] = cb[d];
		}
	}
	int N = (int)tb.parent.size();
	w.box.resize((size_t)N * 6);
	w.lr.resize((size_t)N * 2);
	for (int n = 0; n < N; n++) {
		if (tb.leaf[n] >= 0) {
			for (int d = 0; d < 6; d++) w.box[6 * n + d] = w.tri[(size_t)tb.leaf[n] * MT_REC + 30 + d];
			w.lr[2 * n] = -1;
			w.lr[2 * n + 1] = tb.leaf[n];
		} else {
			for (int d = 0; d < 6; d++) w.box[6 * n + d] = tb.box[6 * n + d];
			w.lr[2 * n] = tb.left[n];
			w.lr[2 * n + 1] = tb.right[n];
		}
	}
	w.root = tb.root;
	// depth first from the root: a query's stack never holds more than depth + 1 nodes
	std::vector<std::pair<int, int>> stk;
	if (tb.root >= 0) stk.push_back({tb.root, 0});
	while (!stk.empty()) {
		auto e = stk.back();
		stk.pop_back();
		if (e.second > w.depth) w.depth = e.second;
		if (tb.leaf[e.first] < 0) {
			stk.push_back({tb.left[e.first], e.second + 1});
			stk.push_back({tb.right[e.first], e.second + 1});
		}
	}
	return w;
}

}  // namespace cpx
