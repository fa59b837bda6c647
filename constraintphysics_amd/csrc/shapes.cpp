// Host-side shape construction.  Follows the reference's Polyhedron
// constructor so that shapes built here carry the same surfaces, winding,
// normal-evaluation vertices, edges and Gauss map (arcs between faces that
// share an edge, antipodal-face redundancy flags, cached self-extrema).
#include "shapes.h"

#include <cmath>
#include <limits>

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

void polyhedron_mass_properties(const std::vector<double>& pts, const std::vector<std::vector<int>>& faces, double density,
                                double* mass, double* tensor9, double* com3) {
	double m = 0, c[3] = {0, 0, 0}, C[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
	// orient faces outward with respect to the vertex centroid
	int np = (int)pts.size() / 3;
	V3 ctr{0, 0, 0};
	for (int i = 0; i < np; i++) ctr += P(pts, i);
	ctr = ctr * (1.0 / np);
	for (const auto& f : faces) {
		V3 a0 = P(pts, f[0]);
		V3 fn{0, 0, 0};
		for (size_t k = 1; k + 1 < f.size(); k++) fn += cross(P(pts, f[k]) - a0, P(pts, f[k + 1]) - a0);
		double sgn = dot(fn, a0 - ctr) < 0 ? -1.0 : 1.0;
		for (size_t k = 1; k + 1 < f.size(); k++) {
			V3 a = a0 - ctr, b = P(pts, f[k]) - ctr, cc = P(pts, f[k + 1]) - ctr;
			double v = sgn * dot(a, cross(b, cc)) / 6.0;
			double dm = density * v;
			m += dm;
			V3 s = a + b + cc;
			c[0] += dm * s.x / 4.0; c[1] += dm * s.y / 4.0; c[2] += dm * s.z / 4.0;
			double A[3] = {a.x, a.y, a.z}, B[3] = {b.x, b.y, b.z}, Cc[3] = {cc.x, cc.y, cc.z}, S[3] = {s.x, s.y, s.z};
			for (int i = 0; i < 3; i++)
				for (int j = 0; j < 3; j++) C[i][j] += dm / 20.0 * (A[i] * A[j] + B[i] * B[j] + Cc[i] * Cc[j] + S[i] * S[j]);
		}
	}
	for (int i = 0; i < 3; i++) c[i] /= m;
	double tr = C[0][0] + C[1][1] + C[2][2];
	double I[3][3];
	for (int i = 0; i < 3; i++)
		for (int j = 0; j < 3; j++) I[i][j] = (i == j ? tr : 0.0) - C[i][j];
	// shift from the centroid to the centre of mass
	double d2 = c[0] * c[0] + c[1] * c[1] + c[2] * c[2];
	for (int i = 0; i < 3; i++)
		for (int j = 0; j < 3; j++) I[i][j] -= m * ((i == j ? d2 : 0.0) - c[i] * c[j]);
	*mass = m;
	for (int i = 0; i < 3; i++)
		for (int j = 0; j < 3; j++) tensor9[3 * i + j] = I[i][j];
	com3[0] = c[0] + ctr.x;
	com3[1] = c[1] + ctr.y;
	com3[2] = c[2] + ctr.z;
}

}  // namespace cpx
