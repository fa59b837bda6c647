// Narrowphase: one 64-lane wavefront per candidate pair.
//
// Convex hull vs convex hull follows SAT_PolyPoly (CollisionDetect.cpp:717-835):
// the candidate axes are A's non-redundant face normals (A's own extrema come
// from the cached self-query, recentred), B's, then the Gauss-map arc pairs
// that pass the hemisphere (:760) and mutual-crossing (:767) tests, with
// n = (a1 x a2) x (b1 x b2).  Lanes evaluate axes in parallel; the reference's
// sequential rule "first strict minimum wins, any separating axis ends the
// test" is reproduced exactly by a wave ballot for separation and a
// lexicographic (depth, axis index) minimum across lanes.
//
// The contact feature search of a hull (CollisionDetect.cpp:237-280) runs on
// the whole wave (adjacent faces and edges tested side by side, a first-minimum
// reduction picks the reference's feature).  The 2-D Sutherland-Hodgman
// clipping with feature-id (MagicID) bookkeeping, lift to 3-D and cull to four
// points (:392-620, :809-831, :2795-2830) are inherently serial and short;
// lane 0 runs them on polygons staged in LDS.
//
// Roofline: per pair the algorithmic traffic is 2 poses + pair key + the
// manifold written (~150 B), but hull vertices are re-read once per tested
// axis, so this kernel is VALU/latency bound rather than HBM bound.
#include "cp_internal.h"

namespace cpx {

#define MAXP 40
#define WAVES_PER_BLOCK 4

struct CEP {
	double x, y;
	uint64_t magic;
	int src;
	int skip;
};

struct Hull {
	const double* wp;  // world points (xyz)
	const double* wn;  // world Gauss normals (xyz)
	int np, nf, na, ne;
	const int *fv_off, *fv;  // indexed by global face id (f0 + local)
	int f0, p0;
	const int *ncalc, *ndir, *sid, *red, *minpid, *maxpid, *refidx;
	const double *gmin, *gmax, *grefval;
	const int *arcs, *edges;
	const int *vaf_off, *vaf, *vae_off, *vae;  // indexed by global vertex id (p0 + local)
};

__device__ __forceinline__ V3 ldv(const double* a, int k) { return V3{a[3 * k], a[3 * k + 1], a[3 * k + 2]}; }

__device__ Hull make_hull(const Bodies& b, const Shapes& sh, int body) {
	Hull h;
	int s = b.shape[body];
	h.wp = b.wpts + 3 * b.wpt_off[body];
	h.wn = b.wnorm + 3 * b.wgv_off[body];
	h.p0 = sh.pt_off[s];
	h.np = sh.pt_off[s + 1] - h.p0;
	h.f0 = sh.face_off[s];
	h.nf = sh.face_off[s + 1] - h.f0;
	int a0 = sh.arc_off[s], e0 = sh.edge_off[s];
	h.na = sh.arc_off[s + 1] - a0;
	h.ne = sh.edge_off[s + 1] - e0;
	h.fv_off = sh.fv_off;
	h.fv = sh.fv;
	h.ncalc = sh.f_ncalc + h.f0;
	h.ndir = sh.f_ndir + h.f0;
	h.sid = sh.f_sid + h.f0;
	h.red = sh.g_red + h.f0;
	h.minpid = sh.g_minpid + h.f0;
	h.maxpid = sh.g_maxpid + h.f0;
	h.refidx = sh.g_refidx + h.f0;
	h.gmin = sh.g_min + h.f0;
	h.gmax = sh.g_max + h.f0;
	h.grefval = sh.g_refval + h.f0;
	h.arcs = sh.arcs + 2 * a0;
	h.edges = sh.edges + 2 * e0;
	h.vaf_off = sh.va_face_off;
	h.vaf = sh.va_faces;
	h.vae_off = sh.va_edge_off;
	h.vae = sh.va_edges;
	return h;
}

// The same hull from the body's geometry record (Bodies::geo, three
// independent 16-byte loads instead of the chain body -> shape -> offsets).
__device__ Hull make_hull_g(const Bodies& b, const Shapes& sh, int body) {
	const int4* g = b.geo + 3 * (size_t)body;
	const int4 g0 = g[0], g1 = g[1], g2 = g[2];
	Hull h;
	h.wp = b.wpts + 3 * (size_t)g0.z;
	h.wn = b.wnorm + 3 * (size_t)g0.w;
	h.p0 = g1.x;
	h.np = g1.y;
	h.f0 = g1.z;
	h.nf = g1.w;
	h.na = g2.y;
	h.ne = g2.w;
	h.fv_off = sh.fv_off;
	h.fv = sh.fv;
	h.ncalc = sh.f_ncalc + h.f0;
	h.ndir = sh.f_ndir + h.f0;
	h.sid = sh.f_sid + h.f0;
	h.red = sh.g_red + h.f0;
	h.minpid = sh.g_minpid + h.f0;
	h.maxpid = sh.g_maxpid + h.f0;
	h.refidx = sh.g_refidx + h.f0;
	h.gmin = sh.g_min + h.f0;
	h.gmax = sh.g_max + h.f0;
	h.grefval = sh.g_refval + h.f0;
	h.arcs = sh.arcs + 2 * g2.x;
	h.edges = sh.edges + 2 * g2.z;
	h.vaf_off = sh.va_face_off;
	h.vaf = sh.va_faces;
	h.vae_off = sh.va_edge_off;
	h.vae = sh.va_edges;
	return h;
}

// Geometry records of the bodies the step visits.
__global__ void k_body_geo(Bodies b, Shapes sh) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= b.nl) return;
	const int s = b.shape[i];
	const int p0 = sh.pt_off[s], f0 = sh.face_off[s], a0 = sh.arc_off[s], e0 = sh.edge_off[s];
	int4* g = b.geo + 3 * (size_t)i;
	g[0] = int4{sh.type[s], s, b.wpt_off[i], b.wgv_off[i]};
	g[1] = int4{p0, sh.pt_off[s + 1] - p0, f0, sh.face_off[s + 1] - f0};
	g[2] = int4{a0, sh.arc_off[s + 1] - a0, e0, sh.edge_off[s + 1] - e0};
}

struct Ext {
	int min_pid, max_pid;
	double min_val, max_val;
};

// The points are read four at a time so that their loads are in flight
// together; the comparisons still run in point order (first extreme wins).
__device__ Ext find_extrema(const Hull& h, V3 axis) {  // CollisionDetect.cpp:627-644
	Ext e{-1, -1, __builtin_inf(), -__builtin_inf()};
	auto take = [&](double v, int i) {
		if (v < e.min_val) { e.min_pid = i; e.min_val = v; }
		if (v > e.max_val) { e.max_pid = i; e.max_val = v; }
	};
	int i = 0;
	for (; i + 4 <= h.np; i += 4) {
		V3 p0 = ldv(h.wp, i), p1 = ldv(h.wp, i + 1), p2 = ldv(h.wp, i + 2), p3 = ldv(h.wp, i + 3);
		double v0 = dot(p0, axis), v1 = dot(p1, axis), v2 = dot(p2, axis), v3 = dot(p3, axis);
		take(v0, i);
		take(v1, i + 1);
		take(v2, i + 2);
		take(v3, i + 3);
	}
	for (; i < h.np; i++) take(dot(ldv(h.wp, i), axis), i);
	return e;
}

__device__ Ext recentered(const Hull& h, int g) {  // CollisionDetect.cpp:622-625 and :726
	double nv = dot(ldv(h.wn, g), ldv(h.wp, h.refidx[g]));
	double diff = nv - h.grefval[g];
	return Ext{h.minpid[g], h.maxpid[g], h.gmin[g] + diff, h.gmax[g] + diff};
}

struct Check {
	int a_pid, b_pid;
	V3 n;
	double pen;
};
__device__ __forceinline__ Check sat_check(const Ext& a, const Ext& b, V3 n) {  // CollisionDetect.cpp:705-715
	double f = a.max_val - b.min_val, r = b.max_val - a.min_val;
	if (f < r) return Check{a.max_pid, b.min_pid, n, f};
	return Check{a.min_pid, b.max_pid, -n, r};
}

// Axis t of the pair's sequential candidate order; valid=false when the arc
// pair is pruned.  Mirrors the three loops of SAT_PolyPoly.
__device__ Check eval_axis(const Hull& A, const Hull& B, int t, bool* valid) {
	*valid = true;
	if (t < A.nf) {
		if (A.red[t]) { *valid = false; return Check{}; }
		V3 n = ldv(A.wn, t);
		return sat_check(recentered(A, t), find_extrema(B, n), n);
	}
	t -= A.nf;
	if (t < B.nf) {
		if (B.red[t]) { *valid = false; return Check{}; }
		V3 n = ldv(B.wn, t);
		return sat_check(find_extrema(A, n), recentered(B, t), n);
	}
	t -= B.nf;
	int i1 = t / B.na, i2 = t - i1 * B.na;
	V3 a1 = ldv(A.wn, A.arcs[2 * i1]), a2 = ldv(A.wn, A.arcs[2 * i1 + 1]);
	V3 b1 = -ldv(B.wn, B.arcs[2 * i2]), b2 = -ldv(B.wn, B.arcs[2 * i2 + 1]);
	V3 aa = a1 + a2;
	if (dot(aa, b1) + dot(aa, b2) <= 0) { *valid = false; return Check{}; }
	V3 ap = cross(a1, a2), bp = cross(b1, b2);
	if (dot(ap, b1) * dot(ap, b2) > 0 || dot(bp, a1) * dot(bp, a2) > 0) { *valid = false; return Check{}; }
	V3 n = cross(ap, bp);
	if (mag2(n) == 0) { *valid = false; return Check{}; }
	n = normalize(n);
	if (dot(aa, n) < 0) n = n * -1.0;
	return sat_check(find_extrema(A, n), find_extrema(B, n), n);
}

// ------------------------------------------------------------- contact area
__device__ V3 face_normal(const Hull& h, int f) {  // Surface::normal, ConvexPrimitive.cpp:815-821
	int o = h.fv_off[h.f0 + f];
	V3 p1 = ldv(h.wp, h.fv[o]), p2 = ldv(h.wp, h.fv[o + 1]), p3 = ldv(h.wp, h.fv[o + h.ncalc[f]]);
	return normalize(cross(p2 - p1, p3 - p2)) * (double)h.ndir[f];
}

// findContactArea (CollisionDetect.cpp:237-280): fills out[].x/y/src, returns count; *surface = id or -1.
__device__ int contact_area(const Hull& h, V3 n, V3 p, int pid, V3 u, V3 w, double cos_tol, double sin_tol, CEP* out, int* surface) {
	int gv = h.p0 + pid;
	int best_f = -1;
	double best_ft = 1.0;
	for (int k = h.vaf_off[gv]; k < h.vaf_off[gv + 1]; k++) {
		int f = h.vaf[k];
		double c = dot(face_normal(h, f), n);
		if (1 - c < best_ft) { best_ft = 1 - c; best_f = f; }
	}
	if (best_ft <= cos_tol) {
		int o = h.fv_off[h.f0 + best_f], e = h.fv_off[h.f0 + best_f + 1];
		int cnt = 0;
		for (int k = o; k < e && cnt < MAXP; k++, cnt++) {
			int vi = h.fv[k];
			V3 v = ldv(h.wp, vi);
			out[cnt].x = dot(v, u);
			out[cnt].y = dot(v, w);
			out[cnt].src = vi;
		}
		*surface = h.sid[best_f];
		return cnt;
	}
	int best_e = -1;
	double best_et = 1.0;
	for (int k = h.vae_off[gv]; k < h.vae_off[gv + 1]; k++) {
		int e = h.vae[k];
		V3 e1 = ldv(h.wp, h.edges[2 * e]), e2 = ldv(h.wp, h.edges[2 * e + 1]);
		double s = fabs(dot(normalize(e2 - e1), n));
		if (s < best_et) { best_et = s; best_e = e; }
	}
	*surface = -1;
	if (best_et <= sin_tol) {
		int i1 = h.edges[2 * best_e], i2 = h.edges[2 * best_e + 1];
		V3 e1 = ldv(h.wp, i1), e2 = ldv(h.wp, i2);
		out[0].x = dot(e1, u); out[0].y = dot(e1, w); out[0].src = i1;
		out[1].x = dot(e2, u); out[1].y = dot(e2, w); out[1].src = i2;
		return 2;
	}
	out[0].x = dot(p, u);
	out[0].y = dot(p, w);
	out[0].src = pid;
	return 1;
}

__device__ __forceinline__ void wave_fence() {
	__builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
	__builtin_amdgcn_wave_barrier();
}

// (value, position) minimum across the wave; every lane ends with the result.
__device__ __forceinline__ void wave_first_min(double* v, int* k) {
	for (int off = 32; off >= 1; off >>= 1) {
		double ov = __shfl_xor(*v, off);
		int ok = __shfl_xor(*k, off);
		if (ov < *v || (ov == *v && ok < *k)) { *v = ov; *k = ok; }
	}
}

// contact_area with the whole wave: every lane calls it with the same
// arguments and gets the same count and surface; out[] is complete after a
// wave_fence.  Lanes test the vertex's adjacent faces (then edges) side by
// side, each keeping its first strict minimum over its own positions; the
// (value, position) minimum across lanes is then the feature the sequential
// loop picks (first strict minimum below 1, NaNs never taken).  The chosen
// face's points are projected one per lane.
__device__ int contact_area_wave(const Hull& h, V3 n, V3 p, int pid, V3 u, V3 w, double cos_tol, double sin_tol, CEP* out, int* surface,
                                 int lane) {
	const int gv = h.p0 + pid;
	double bt = 1.0;
	int bk = 0x7FFFFFFF;
	for (int k = h.vaf_off[gv] + lane; k < h.vaf_off[gv + 1]; k += 64) {
		double t = 1 - dot(face_normal(h, h.vaf[k]), n);
		if (t < bt) { bt = t; bk = k; }
	}
	wave_first_min(&bt, &bk);
	if (bk != 0x7FFFFFFF && bt <= cos_tol) {
		int f = h.vaf[bk];
		int o = h.fv_off[h.f0 + f], e = h.fv_off[h.f0 + f + 1];
		int cnt = e - o < MAXP ? e - o : MAXP;
		for (int k = lane; k < cnt; k += 64) {
			int vi = h.fv[o + k];
			V3 v = ldv(h.wp, vi);
			out[k].x = dot(v, u);
			out[k].y = dot(v, w);
			out[k].src = vi;
		}
		*surface = h.sid[f];
		return cnt;
	}
	double be = 1.0;
	int ke = 0x7FFFFFFF;
	for (int k = h.vae_off[gv] + lane; k < h.vae_off[gv + 1]; k += 64) {
		int e = h.vae[k];
		V3 e1 = ldv(h.wp, h.edges[2 * e]), e2 = ldv(h.wp, h.edges[2 * e + 1]);
		double s = fabs(dot(normalize(e2 - e1), n));
		if (s < be) { be = s; ke = k; }
	}
	wave_first_min(&be, &ke);
	*surface = -1;
	if (ke != 0x7FFFFFFF && be <= sin_tol) {
		if (lane < 2) {
			int vi = h.edges[2 * h.vae[ke] + lane];
			V3 v = ldv(h.wp, vi);
			out[lane].x = dot(v, u);
			out[lane].y = dot(v, w);
			out[lane].src = vi;
		}
		return 2;
	}
	if (lane == 0) {
		out[0].x = dot(p, u);
		out[0].y = dot(p, w);
		out[0].src = pid;
	}
	return 1;
}

// ------------------------------------------------------------- clipping
__device__ __forceinline__ uint32_t edge_id(uint16_t a, uint16_t b) {  // CollisionDetect.cpp:9-21
	int mn, mx;
	if (a < b) { mn = a; mx = b; } else { mn = b; mx = a; }
	return (0x0000FFFFu & (uint32_t)mn) + (0xFFFF0000u & ((uint32_t)mx << 16));
}
__device__ __forceinline__ double dot2(double ax, double ay, double bx, double by) {
	double s = 0;
	s += ax * bx;
	s += ay * by;
	return s;
}
__device__ __forceinline__ void in_dir(double p1x, double p1y, double p2x, double p2y, double* ox, double* oy) {  // :392-396
	double dx = p1y - p2y, dy = p2x - p1x;
	double s = 0;
	s += dx * dx;
	s += dy * dy;
	double k = 1.0 / sqrt(s);
	*ox = dx * k;
	*oy = dy * k;
}
__device__ bool winding_ccw(const CEP* c, int n) {  // :398-409
	double ix, iy;
	in_dir(c[0].x, c[0].y, c[1].x, c[1].y, &ix, &iy);
	for (int i = 2; i < n; i++) {
		double d = dot2(c[i].x - c[0].x, c[i].y - c[0].y, ix, iy);
		if (d > 0.0000000001) return true;
		else if (d < -0.0000000001) return false;
	}
	return false;
}
// createClipEvaluationPoly (:423-442), in place.
__device__ void clip_poly(CEP* c, int n, uint32_t other_sid, bool flip) {
	for (int i = 0; i < n; i++) {
		uint64_t m = 0;
		if (flip) {
			m |= 0x00000000FFFFFFFFull & (uint64_t)other_sid;
			m |= 0xFFFFFFFF00000000ull & ((uint64_t)(int64_t)c[i].src << 32);
		} else {
			m |= 0x00000000FFFFFFFFull & (uint64_t)(int64_t)c[i].src;
			m |= 0xFFFFFFFF00000000ull & ((uint64_t)other_sid << 32);
		}
		c[i].magic = m;
		c[i].skip = 0;
	}
	if (n > 2 && !winding_ccw(c, n)) {
		for (int i = 0; i < n / 2; i++) {
			CEP t = c[i];
			c[i] = c[n - 1 - i];
			c[n - 1 - i] = t;
		}
	}
}
// getEdgeIntersectionWithClippingEdge (:444-467).
__device__ CEP intersect(const CEP& e1, const CEP& e2, double cnx, double cny, double spx, double spy, int32_t ceid, bool flip) {
	double dx = e2.x - e1.x, dy = e2.y - e1.y;
	double t = dot2(cnx, cny, spx - e1.x, spy - e1.y) / dot2(cnx, cny, dx, dy);
	CEP o;
	o.x = e1.x + dx * t;
	o.y = e1.y + dy * t;
	bool e2p = dot2(dx, dy, cnx, cny) > 0;
	o.src = e2p ? e1.src : e2.src;
	uint64_t m = 0;
	uint32_t eid = edge_id((uint16_t)e1.src, (uint16_t)e2.src);
	if (flip) {
		m |= 0x00000000FFFFFFFFull & (uint64_t)(int64_t)ceid;
		m |= 0xFFFFFFFF00000000ull & ((uint64_t)eid << 32);
	} else {
		m |= 0x00000000FFFFFFFFull & (uint64_t)eid;
		m |= 0xFFFFFFFF00000000ull & ((uint64_t)(int64_t)ceid << 32);
	}
	o.magic = m;
	o.skip = !e2p;
	return o;
}
enum { BOTH, P1ON_P2C, P2ON_P1C, NEITHER, ONLY_P1, ONLY_P2 };
__device__ __forceinline__ int check_seg(double a, double b, double c) {  // :473-480
	if (a >= c && b >= c) return NEITHER;
	if (a == c && b < c) return P1ON_P2C;
	if (a < c && b == c) return P2ON_P1C;
	if (a < c && b >= c) return ONLY_P1;
	if (a >= c && b < c) return ONLY_P2;
	return BOTH;
}
// getClipEvaluatinPolyAfterClippingByEdge (:482-562): src[0..n) -> dst, returns count.
__device__ int clip_by_edge(const CEP* c, int n, double sx, double sy, double spx, double spy, int32_t eid, bool flip, CEP* out) {
	double cv = dot2(spx, spy, sx, sy);
	int k = 0;
	if (n >= 3) {
		double p1v = 0, p2v = 0;
		CEP p1, p2;
		for (int i = 0; i < n; i++) {
			if (i > 0) { p1v = p2v; p1 = p2; }
			else { p1 = c[0]; p1v = dot2(p1.x, p1.y, sx, sy); }
			p2 = c[(i + 1) % n];
			p2v = dot2(p2.x, p2.y, sx, sy);
			int r = check_seg(p1v, p2v, cv);
			if (r == NEITHER || r == P1ON_P2C) {
				if (k < MAXP) out[k++] = p1;
			} else if (r == ONLY_P2) {
				if (k < MAXP) out[k++] = p1;
				if (k < MAXP) out[k++] = intersect(p1, p2, sx, sy, spx, spy, eid, flip);
			} else if (r == ONLY_P1) {
				if (k < MAXP) out[k++] = intersect(p1, p2, sx, sy, spx, spy, eid, flip);
			}
		}
		return k;
	} else if (n == 2) {
		double a = dot2(c[0].x, c[0].y, sx, sy), b = dot2(c[1].x, c[1].y, sx, sy);
		int r = check_seg(a, b, cv);
		if (r == NEITHER) { out[0] = c[0]; out[1] = c[1]; return 2; }
		if (r == P1ON_P2C) { out[0] = c[0]; return 1; }
		if (r == P2ON_P1C) { out[0] = c[1]; return 1; }
		if (r == ONLY_P1) { CEP t = intersect(c[0], c[1], sx, sy, spx, spy, eid, flip); out[1] = c[1]; out[0] = t; return 2; }
		if (r == ONLY_P2) { CEP t = intersect(c[0], c[1], sx, sy, spx, spy, eid, flip); out[0] = c[0]; out[1] = t; return 2; }
		return 0;
	} else if (n == 1) {
		if (dot2(c[0].x, c[0].y, sx, sy) >= cv) { out[0] = c[0]; return 1; }
		return 0;
	}
	return 0;
}
// clipC1ByAllEdgesOfC2 (:564-576).  c1 lives in bufA (scratch bufB); returns
// the buffer holding the result and its count.
__device__ CEP* clip_all(CEP* bufA, CEP* bufB, int n1, const CEP* c2, int n2, bool flip, int* nout) {
	CEP* cur = bufA;
	CEP* nxt = bufB;
	for (int i = 0; i < n2; i++) {
		CEP e1 = c2[i];
		if (e1.skip) continue;
		CEP e2 = c2[(i + 1) % n2];
		double dx, dy;
		in_dir(e1.x, e1.y, e2.x, e2.y, &dx, &dy);
		int32_t eid = (int32_t)edge_id((uint16_t)e1.src, (uint16_t)e2.src);
		n1 = clip_by_edge(cur, n1, dx, dy, e1.x, e1.y, eid, flip, nxt);
		CEP* t = cur;
		cur = nxt;
		nxt = t;
		if (n1 == 0) break;
	}
	*nout = n1;
	return cur;
}

__device__ __forceinline__ uint64_t swap_halves(uint64_t v) {  // swapOrder (CollisionDetect.cpp:28-40)
	uint64_t lo = v & 0x00000000FFFFFFFFull, hi = (v & 0xFFFFFFFF00000000ull) >> 32;
	return (0x00000000FFFFFFFFull & hi) + (0xFFFFFFFF00000000ull & (lo << 32));
}

// clipContacts (CollisionDetect.cpp:578-620) on staged areas: b0 = lds[0..)
// holds A's na points (surface sa), b1 = lds[MAXP..) B's nb points (surface
// sb); the rest of lds is scratch.  Returns the buffer holding the pool.
__device__ CEP* clip_areas(CEP* lds, int na, int sa, int nb, int sb, int* nr) {
	CEP* b0 = lds;
	CEP* b1 = lds + MAXP;
	CEP* b2 = lds + 2 * MAXP;
	clip_poly(b0, na, (uint32_t)sb, false);
	clip_poly(b1, nb, (uint32_t)sa, true);
	if (na == 1) { *nr = 1; return b0; }
	if (nb == 1) { *nr = 1; return b1; }
	if (na == 2 && nb == 2) {
		double dx, dy;
		in_dir(b1[0].x, b1[0].y, b1[1].x, b1[1].y, &dx, &dy);
		b2[0] = intersect(b0[0], b0[1], dx, dy, b1[0].x, b1[0].y, (int32_t)edge_id((uint16_t)b1[0].src, (uint16_t)b1[1].src), false);
		*nr = 1;
		return b2;
	}
	if (na == 2) return clip_all(b0, b2, na, b1, nb, false, nr);
	if (nb == 2) return clip_all(b1, b2, nb, b0, na, true, nr);
	// poly v poly: clip A by B's edges, then B by the clipped A's edges
	// (an empty clipped A leaves B unclipped, as in the reference).
	int n1;
	CEP* r1 = clip_all(b0, b2, na, b1, nb, false, &n1);
	CEP* scratch = (r1 == b0) ? b2 : b0;
	return clip_all(b1, scratch, nb, r1, n1, true, nr);
}

// Lift the pool to 3-D (e.g. CollisionDetect.cpp:809-831) and write the slot,
// culled to four points (cull_manifold, :2795-2830, with the reference's own
// cos/sin of 2*pi*i/4).  n is the SAT normal the pool was projected with;
// flip applies detectCollision's swap for reversed pairs (normal negated,
// MagicID halves swapped) before the cull, which works on the final normal.
__device__ void emit_manifold(const CEP* res, int nr, V3 n, V3 u, V3 w, V3 amax, double pen, uint64_t cid, bool flip, const Params& P,
                              int slot, Manifolds& M, int cap) {
	double a_dot = dot(amax, n);
	V3 noff = n * a_dot;
	V3 nf = flip ? -n : n;
	if (flip) cid = swap_halves(cid);
	M.normal[slot] = nf.x;
	M.normal[cap + slot] = nf.y;
	M.normal[2 * cap + slot] = nf.z;
	double* mpts = M.pts + (size_t)slot * 16;
	uint64_t* mmag = M.magic + (size_t)slot * 8;
	if (nr <= 4) {
		for (int k = 0; k < nr; k++) {
			V3 pos = u * res[k].x + w * res[k].y + noff;
			mpts[4 * k] = pos.x; mpts[4 * k + 1] = pos.y; mpts[4 * k + 2] = pos.z;
			mpts[4 * k + 3] = dot(pos, n) - a_dot + pen;
			mmag[2 * k] = cid;
			mmag[2 * k + 1] = flip ? swap_halves(res[k].magic) : res[k].magic;
		}
		M.count[slot] = nr;
		M.maxpen[slot] = pen;
		return;
	}
	uint64_t avail = (nr >= 64) ? ~0ull : ((1ull << nr) - 1);
	V3 cu, cw;
	perp_basis(nf, &cu, &cw);
	double maxpen = 0;
	for (int i = 0; i < 4; i++) {
		V3 t = cu * P.cull_c[i] + cw * P.cull_s[i];
		double bv = -__builtin_inf();
		int bi = -1;
		for (int j = 0; j < nr; j++) {
			if (!((avail >> j) & 1)) continue;
			V3 pos = u * res[j].x + w * res[j].y + noff;
			double v = dot(pos, t);
			if (v > bv) { bv = v; bi = j; }
		}
		if (bi < 0) bi = 0;
		V3 pos = u * res[bi].x + w * res[bi].y + noff;
		double pp = dot(pos, n) - a_dot + pen;
		mpts[4 * i] = pos.x; mpts[4 * i + 1] = pos.y; mpts[4 * i + 2] = pos.z; mpts[4 * i + 3] = pp;
		mmag[2 * i] = cid;
		mmag[2 * i + 1] = flip ? swap_halves(res[bi].magic) : res[bi].magic;
		maxpen = (maxpen < pp) ? pp : maxpen;
		avail &= ~(1ull << bi);
	}
	M.count[slot] = 4;
	M.maxpen[slot] = maxpen;
}

// Contact generation for a hull pair after SAT, called by the whole wave with
// the same mp: the two contact areas are found by the wave, the clipping and
// the manifold write run on lane 0.
__device__ void hull_contacts(const Hull& A, const Hull& B, const Check& mp, const Params& P, uint64_t cid, CEP* lds, int slot,
                              Manifolds& M, int cap, int lane) {
	V3 n = mp.n;
	V3 amax = ldv(A.wp, mp.a_pid), bmax = ldv(B.wp, mp.b_pid);
	V3 u, w;
	perp_basis(n, &u, &w);
	int sa, sb;
	int na = contact_area_wave(A, n, amax, mp.a_pid, u, w, P.cos_tol, P.sin_tol, lds, &sa, lane);
	int nb = contact_area_wave(B, n * -1.0, bmax, mp.b_pid, u, w, P.cos_tol, P.sin_tol, lds + MAXP, &sb, lane);
	wave_fence();
	if (lane == 0) {
		int nr;
		CEP* res = clip_areas(lds, na, sa, nb, sb, &nr);
		emit_manifold(res, nr, n, u, w, amax, mp.pen, cid, false, P, slot, M, cap);
	}
}

// ------------------------------------------------------------- spheres
__device__ __forceinline__ Ext sphere_extrema(V3 c, double r, V3 dir) {  // CollisionDetect.cpp:646-656
	double cv = dot(dir, c);
	return Ext{-1, -1, cv - r, cv + r};
}

// Candidate axis t of SAT_PolySphere (CollisionDetect.cpp:1336-1385): A's
// non-redundant face normals, then vertex directions, then edge normals.
__device__ Check eval_axis_sphere(const Hull& A, V3 c, double r, int t, bool* valid, uint32_t* feat) {
	*valid = true;
	if (t < A.nf) {
		if (A.red[t]) { *valid = false; return Check{}; }
		V3 n = ldv(A.wn, t);
		*feat = (uint32_t)(A.np + t);
		return sat_check(recentered(A, t), sphere_extrema(c, r, n), n);
	}
	t -= A.nf;
	if (t < A.np) {
		V3 n = normalize(ldv(A.wp, t) - c);
		*feat = (uint32_t)t;
		return sat_check(find_extrema(A, n), sphere_extrema(c, r, n), n);
	}
	t -= A.np;
	int i1 = A.edges[2 * t], i2 = A.edges[2 * t + 1];
	V3 e1 = ldv(A.wp, i1), e2 = ldv(A.wp, i2);
	V3 dir = normalize(e2 - e1);
	V3 sample = e1 - c;
	V3 n = normalize(sample - dir * dot(dir, sample));
	*feat = edge_id((uint16_t)i1, (uint16_t)i2);
	return sat_check(find_extrema(A, n), sphere_extrema(c, r, n), n);
}

__device__ void write_mat(const Bodies& b, int ia, int ib, Manifolds& M, int slot, int cap) {  // CollisionDetect.cpp:820-824
	double ra = b.mat[ia], rb = b.mat[ib];
	double r = (ra < rb) ? rb : ra;
	double mk = (b.mat[b.n + ia] + b.mat[b.n + ib]) / 2.0;
	double ms = (b.mat[2 * b.n + ia] + b.mat[2 * b.n + ib]) / 2.0;
	double* o = M.mat + (size_t)slot * 12;
	for (int k = 0; k < 4; k++) {
		o[3 * k] = r;
		o[3 * k + 1] = mk;
		o[3 * k + 2] = ms;
	}
}

// Wave-parallel SAT_PolySphere; poly = body ip, sphere = body is_.  When the
// sphere is the pair's first body the manifold is flipped as detectCollision
// does (CollisionDetect.cpp:91-99): normal negated, MagicID halves swapped.
__device__ void poly_sphere(const Bodies& b, const Shapes& sh, const Hull& A, int ip, int is_, bool swapped, int lane, int slot, Manifolds& M,
                            int cap) {
	V3 c = ldv(b.wpts + 3 * b.wpt_off[is_], 0);
	double r = sh.radius[b.shape[is_]];
	int naxes = A.nf + A.np + A.ne;
	double bp = __builtin_inf();
	int bi = 0x7FFFFFFF;
	int sep = 0;
	for (int base = 0; base < naxes; base += 64) {
		int t = base + lane;
		if (t < naxes) {
			bool valid;
			uint32_t f;
			Check x = eval_axis_sphere(A, c, r, t, &valid, &f);
			if (valid) {
				if (x.pen < 0) sep = 1;
				if (x.pen < bp) { bp = x.pen; bi = t; }
			}
		}
		if (__ballot(sep) != 0) break;
	}
	if (__ballot(sep) != 0) return;
	for (int off = 32; off >= 1; off >>= 1) {
		double obp = __shfl_xor(bp, off);
		int obi = __shfl_xor(bi, off);
		if (obp < bp || (obp == bp && obi < bi)) { bp = obp; bi = obi; }
	}
	if (!(bp > 0) || bi == 0x7FFFFFFF || lane != 0) return;
	bool valid;
	uint32_t feat = 0;
	Check mp = eval_axis_sphere(A, c, r, bi, &valid, &feat);
	uint64_t kp = (uint64_t)(ip + 1) * 16, ks = (uint64_t)(is_ + 1) * 16;
	uint64_t cid = (0x00000000FFFFFFFFull & kp) | (0xFFFFFFFF00000000ull & (ks << 32));
	uint64_t bid = (uint64_t)feat;
	V3 n = mp.n;
	V3 pos = c - n * r;
	if (swapped) {
		n = -n;
		cid = swap_halves(cid);
		bid = swap_halves(bid);
	}
	M.normal[slot] = n.x; M.normal[cap + slot] = n.y; M.normal[2 * cap + slot] = n.z;
	double* mpts = M.pts + (size_t)slot * 16;
	mpts[0] = pos.x; mpts[1] = pos.y; mpts[2] = pos.z; mpts[3] = mp.pen;
	M.magic[(size_t)slot * 8] = cid;
	M.magic[(size_t)slot * 8 + 1] = bid;
	M.maxpen[slot] = mp.pen;
	M.count[slot] = 1;
	write_mat(b, ip, is_, M, slot, cap);
}

// detectSphereSphere (CollisionDetect.cpp:837-866), lane 0 only.
__device__ void sphere_sphere(const Bodies& b, const Shapes& sh, int ia, int ib, int slot, Manifolds& M, int cap) {
	V3 ca = ldv(b.wpts + 3 * b.wpt_off[ia], 0), cb = ldv(b.wpts + 3 * b.wpt_off[ib], 0);
	double ra = sh.radius[b.shape[ia]], rb = sh.radius[b.shape[ib]];
	V3 diff = cb - ca;
	double d = mag(diff);
	double pen = ra + rb - d;
	if (!(pen > 0)) return;
	V3 n = normalize(diff);
	V3 pos = ca + n * ra;
	uint64_t ka = (uint64_t)(ia + 1) * 16, kb = (uint64_t)(ib + 1) * 16;
	M.normal[slot] = n.x; M.normal[cap + slot] = n.y; M.normal[2 * cap + slot] = n.z;
	double* mpts = M.pts + (size_t)slot * 16;
	mpts[0] = pos.x; mpts[1] = pos.y; mpts[2] = pos.z; mpts[3] = pen;
	M.magic[(size_t)slot * 8] = (0x00000000FFFFFFFFull & ka) | (0xFFFFFFFF00000000ull & (kb << 32));
	M.magic[(size_t)slot * 8 + 1] = 0;
	M.maxpen[slot] = pen;
	M.count[slot] = 1;
	write_mat(b, ia, ib, M, slot, cap);
}


// ------------------------------------------------------------- cylinders
// World cylinder of a body: points [centre, top cap 8-gon, bottom cap 8-gon],
// normals [height axis, 14 Gauss vertices] (k_refresh_geometry writes them).
struct Cyl {
	V3 c, ax;
	double r, h;
	const double *top, *bot, *gv;
};
__device__ Cyl make_cyl_g(const Bodies& b, const Shapes& sh, int body) {
	const int4 g0 = b.geo[3 * (size_t)body];
	const double* wp = b.wpts + 3 * (size_t)g0.z;
	const double* wn = b.wnorm + 3 * (size_t)g0.w;
	Cyl c;
	c.c = ldv(wp, 0);
	c.top = wp + 3;
	c.bot = wp + 27;
	c.ax = ldv(wn, 0);
	c.gv = wn + 3;
	c.r = sh.radius[g0.y];
	c.h = sh.height[g0.y];
	return c;
}
__device__ Cyl make_cyl(const Bodies& b, const Shapes& sh, int body) {
	int s = b.shape[body];
	const double* wp = b.wpts + 3 * b.wpt_off[body];
	const double* wn = b.wnorm + 3 * b.wgv_off[body];
	Cyl c;
	c.c = ldv(wp, 0);
	c.top = wp + 3;
	c.bot = wp + 27;
	c.ax = ldv(wn, 0);
	c.gv = wn + 3;
	c.r = sh.radius[s];
	c.h = sh.height[s];
	return c;
}
// MagicID ids of an 8-gon cylinder (ConvexPrimitive.h:174-179).
#define CYL_TOP_PT 0
#define CYL_BOT_PT 8
#define CYL_TOP_SURF 16
#define CYL_BOT_SURF 17
#define CYL_TOP_EDGE 18
#define CYL_BOT_EDGE 19

__device__ __forceinline__ V3 cyl_top_center(const Cyl& c) { return c.c + c.ax * (0.5 * c.h); }  // ConvexPrimitive.h:182
__device__ __forceinline__ V3 cyl_bot_center(const Cyl& c) { return c.c - c.ax * (0.5 * c.h); }  // :183

// Cylinder::getExtremaOfDisk (ConvexPrimitive.cpp:319-355).
__device__ V3 disk_extrema(V3 dc, V3 dn, double r, V3 dir) {
	double nd = dot(dn, dir);
	if (fabs(nd) > 0.99999999) return dc;
	if (fabs(nd) < 0.00000001) return dc + dir * r;
	V3 u, v;
	V3 w = dir;
	perp_basis(dir, &u, &v);
	double nu = dot(dn, u), nv = dot(dn, v), nw = dot(dn, w);
	double dwdu = -nu / nw;
	double dwdv = -nv / nw;
	double gm = sqrt(dwdu * dwdu + dwdv * dwdv);
	double gu = dwdu / gm;
	double gvv = dwdv / gm;
	double uv2 = r * r / (1 + gm * gm);
	double dw = sqrt(r * r - uv2);
	double du = gu * sqrt(uv2);
	double dv = gvv * sqrt(uv2);
	return dc + u * du + v * dv + w * dw;
}

// getCylinderExtrema (CollisionDetect.cpp:658-684).
__device__ Ext cyl_extrema(const Cyl& c, V3 dir) {
	V3 tc = cyl_top_center(c);
	V3 tmax = disk_extrema(tc, c.ax, c.r, dir);
	V3 tmin = tc * 2.0 - tmax;
	V3 bmax = c.c * 2.0 - tmin;
	V3 bmin = c.c * 2.0 - tmax;
	double tmaxv = dot(dir, tmax), tminv = dot(dir, tmin), bmaxv = dot(dir, bmax), bminv = dot(dir, bmin);
	Ext e;
	e.max_val = (tmaxv < bmaxv) ? bmaxv : tmaxv;  // std::max<double>
	e.min_val = (bminv < tminv) ? bminv : tminv;  // std::min<double>
	e.min_pid = -1;
	e.max_pid = -1;
	return e;
}

// cylinderLengthwiseLineInDirection (CollisionDetect.cpp:347-352).
__device__ void cyl_line(V3 center, V3 ax, double r, double h, V3 n, V3* p1, V3* p2) {
	V3 barrel = normalize(n - ax * dot(ax, n));
	*p1 = center + barrel * r + ax * (0.5 * h);
	*p2 = *p1 - ax * h;
}

enum { AREA_FACE = 0, AREA_EDGE = 1, AREA_VERTEX = 2, AREA_BARREL = 3 };
// findCylinderContactArea (CollisionDetect.cpp:354-374); returns the count,
// *surface and *origin.
__device__ int cyl_contact_area(const Cyl& c, V3 n, V3 u, V3 w, double cos_tol, double sin_tol, CEP* out, int* surface, int* origin) {
	double an = dot(c.ax, n);
	if (1 - fabs(an) <= cos_tol) {
		bool top = dot(n, c.ax) > 0;
		const double* fv = top ? c.top : c.bot;
		for (int i = 0; i < 8; i++) {
			V3 v = ldv(fv, i);
			out[i].x = dot(v, u);
			out[i].y = dot(v, w);
			out[i].src = i + (top ? CYL_TOP_PT : CYL_BOT_PT);
		}
		*surface = top ? CYL_TOP_SURF : CYL_BOT_SURF;
		*origin = AREA_FACE;
		return 8;
	}
	*surface = -1;
	if (fabs(an) <= sin_tol) {
		V3 p1, p2;
		cyl_line(c.c, c.ax, c.r, c.h, n, &p1, &p2);
		out[0].x = dot(p1, u); out[0].y = dot(p1, w); out[0].src = CYL_TOP_EDGE;
		out[1].x = dot(p2, u); out[1].y = dot(p2, w); out[1].src = CYL_BOT_EDGE;
		*origin = AREA_BARREL;
		return 2;
	}
	bool top = dot(n, c.ax) > 0;
	V3 p = disk_extrema(top ? cyl_top_center(c) : cyl_bot_center(c), c.ax, c.r, n);
	out[0].x = dot(p, u); out[0].y = dot(p, w); out[0].src = top ? CYL_TOP_EDGE : CYL_BOT_EDGE;
	*origin = AREA_EDGE;
	return 1;
}

// Gauss-arc pair direction (arc loop body of CollisionDetect.cpp:1617-1650,
// :1021-1054); false when pruned.
// The acceptance tests of arc_axis alone (no axis), written as negations of
// the reference's `continue` conditions so NaNs take the same branch.
__device__ __forceinline__ bool arc_test(V3 a1, V3 a2, V3 aa, V3 ap, V3 b1, V3 b2, V3 bp) {
	if (dot(aa, b1) + dot(aa, b2) <= 0) return false;
	if (dot(ap, b1) * dot(ap, b2) > 0 || dot(bp, a1) * dot(bp, a2) > 0) return false;
	return !(mag2(cross(ap, bp)) == 0);
}
__device__ __forceinline__ bool arc_pair(V3 a1, V3 a2, V3 b1, V3 b2) { return arc_test(a1, a2, a1 + a2, cross(a1, a2), b1, b2, cross(b1, b2)); }
__device__ bool arc_axis(V3 a1, V3 a2, V3 b1, V3 b2, V3* out) {
	V3 aa = a1 + a2;
	if (dot(aa, b1) + dot(aa, b2) <= 0) return false;
	V3 ap = cross(a1, a2), bp = cross(b1, b2);
	if (dot(ap, b1) * dot(ap, b2) > 0 || dot(bp, a1) * dot(bp, a2) > 0) return false;
	V3 n = cross(ap, bp);
	if (mag2(n) == 0) return false;
	n = normalize(n);
	if (dot(aa, n) < 0) n = n * -1.0;
	*out = n;
	return true;
}
// Cylinder Gauss arc k as the reference builds it: {0, k/2} or {1, k/2}
// (ConvexPrimitive.cpp:239-245, second index = loop counter); 24 arcs.
__device__ __forceinline__ V3 cyl_gv(const Cyl& c, int i) { return ldv(c.gv, i); }

// Candidate axis t of SAT_PolyCylinder (CollisionDetect.cpp:1563-1650): poly
// face normals, the cylinder axis, poly edge x axis, vertex-to-axis
// directions, then poly arc x cylinder arc pairs.
__device__ Check eval_axis_pc(const Hull& A, const Cyl& B, int t, bool* valid) {
	*valid = true;
	if (t < A.nf) {
		if (A.red[t]) { *valid = false; return Check{}; }
		V3 n = ldv(A.wn, t);
		return sat_check(recentered(A, t), cyl_extrema(B, n), n);
	}
	t -= A.nf;
	if (t == 0) return sat_check(find_extrema(A, B.ax), cyl_extrema(B, B.ax), B.ax);
	t -= 1;
	if (t < A.ne) {
		V3 ed = ldv(A.wp, A.edges[2 * t + 1]) - ldv(A.wp, A.edges[2 * t]);
		V3 dir = cross(ed, B.ax);
		if (mag(dir) < 0.00000000001) { *valid = false; return Check{}; }
		V3 dn = normalize(dir);
		return sat_check(find_extrema(A, dn), cyl_extrema(B, dn), dn);
	}
	t -= A.ne;
	if (t < A.np) {
		V3 diff = ldv(A.wp, t) - B.c;
		V3 n = normalize(diff - B.ax * dot(B.ax, diff));
		return sat_check(find_extrema(A, n), cyl_extrema(B, n), n);
	}
	t -= A.np;
	int i1 = t / 24, k = t - i1 * 24;
	V3 n;
	if (!arc_axis(ldv(A.wn, A.arcs[2 * i1]), ldv(A.wn, A.arcs[2 * i1 + 1]), -cyl_gv(B, k & 1), -cyl_gv(B, k >> 1), &n)) {
		*valid = false;
		return Check{};
	}
	return sat_check(find_extrema(A, n), cyl_extrema(B, n), n);
}

// Candidate axis t of detectCylinderCylinder (CollisionDetect.cpp:958-1054):
// both axes, cap-polygon edges x the other axis (A top, A bottom, B top, B
// bottom), the barrel-barrel direction, then 24 x 24 arc pairs.
__device__ Check eval_axis_cc(const Cyl& A, const Cyl& B, int t, bool* valid) {
	*valid = true;
	if (t < 2) {
		V3 n = t == 0 ? A.ax : B.ax;
		return sat_check(cyl_extrema(A, n), cyl_extrema(B, n), n);
	}
	t -= 2;
	if (t < 32) {
		int g = t >> 3, i = t & 7;
		const double* poly = g == 0 ? A.top : g == 1 ? A.bot : g == 2 ? B.top : B.bot;
		V3 bax = g < 2 ? B.ax : A.ax;
		V3 ed = ldv(poly, (i + 1) % 8) - ldv(poly, i);
		V3 dir = cross(ed, bax);
		if (mag(dir) < 0.00000000001) { *valid = false; return Check{}; }
		V3 dn = normalize(dir);
		return sat_check(cyl_extrema(A, dn), cyl_extrema(B, dn), dn);
	}
	t -= 32;
	if (t == 0) {
		V3 bb = cross(A.ax, B.ax);
		if (!(mag2(bb) > 0.00000001)) { *valid = false; return Check{}; }
		bb = normalize(bb);
		return sat_check(cyl_extrema(A, bb), cyl_extrema(B, bb), bb);
	}
	t -= 1;
	int k1 = t / 24, k2 = t - k1 * 24;
	V3 n;
	if (!arc_axis(cyl_gv(A, k1 & 1), cyl_gv(A, k1 >> 1), -cyl_gv(B, k2 & 1), -cyl_gv(B, k2 >> 1), &n)) {
		*valid = false;
		return Check{};
	}
	return sat_check(cyl_extrema(A, n), cyl_extrema(B, n), n);
}

__device__ __forceinline__ uint64_t pair_cid(int ia, int ib) {
	uint64_t ka = (uint64_t)(ia + 1) * 16, kb = (uint64_t)(ib + 1) * 16;
	return (0x00000000FFFFFFFFull & ka) | (0xFFFFFFFF00000000ull & (kb << 32));
}

// Contact generation of SAT_PolyCylinder (CollisionDetect.cpp:1652-1709) after
// SAT, called by the whole wave with the same mp: the hull's contact area is
// found by the wave, the rest runs on lane 0; swapped when the cylinder is the
// pair's first body.
__device__ void poly_cyl_contacts(const Hull& A, const Cyl& B, const Check& mp, const Params& P, int ip, int ic, bool swapped, CEP* lds,
                                  int slot, Manifolds& M, int cap, int lane) {
	V3 n = mp.n;
	V3 amax = ldv(A.wp, mp.a_pid);
	V3 u, w;
	perp_basis(n, &u, &w);
	int sa, sb, ob;
	int na = contact_area_wave(A, n, amax, mp.a_pid, u, w, P.cos_tol, P.sin_tol, lds, &sa, lane);
	wave_fence();
	if (lane != 0) return;
	int nb = cyl_contact_area(B, -n, u, w, P.cos_tol, P.sin_tol, lds + MAXP, &sb, &ob);
	CEP* res;
	int nr;
	if (ob == AREA_EDGE) {
		res = lds + MAXP;
		res[0].magic = 0;
		nr = 1;
	} else res = clip_areas(lds, na, sa, nb, sb, &nr);
	emit_manifold(res, nr, n, u, w, amax, mp.pen, pair_cid(ip, ic), swapped, P, slot, M, cap);
}

// Contact generation of detectCylinderCylinder (CollisionDetect.cpp:1056-1137)
// after SAT, lane 0.
__device__ void cyl_cyl_contacts(const Cyl& A, const Cyl& B, const Check& mp, const Params& P, int ia, int ib, CEP* lds, int slot,
                                 Manifolds& M, int cap) {
	V3 n = mp.n;
	V3 u, w;
	perp_basis(n, &u, &w);
	int sa, sb, oa, ob;
	int na = cyl_contact_area(A, n, u, w, P.cos_tol, P.sin_tol, lds, &sa, &oa);
	int nb = cyl_contact_area(B, -n, u, w, P.cos_tol, P.sin_tol, lds + MAXP, &sb, &ob);
	CEP* res;
	int nr;
	if (oa == AREA_EDGE) {
		res = lds;
		res[0].magic = 0;
		nr = 1;
	} else if (ob == AREA_EDGE) {
		res = lds + MAXP;
		res[0].magic = 0;
		nr = 1;
	} else if (oa == AREA_BARREL && ob == AREA_BARREL && fabs(dot(A.ax, B.ax)) > 0.995) {
		// parallel barrels (:1082-1104): overlap of the two lines along A's axis
		V3 a1, a2, b1, b2;
		cyl_line(A.c, A.ax, A.r, A.h, n, &a1, &a2);
		cyl_line(B.c, B.ax, B.r, B.h, -n, &b1, &b2);
		if (dot(B.ax, A.ax) < 0) { V3 t = b1; b1 = b2; b2 = t; }
		double amaxv = dot(a1, A.ax), aminv = dot(a2, A.ax), bmaxv = dot(b1, A.ax), bminv = dot(b2, A.ax);
		V3 imax = (amaxv > bmaxv) ? b1 : a1;
		V3 imin = (aminv > bminv) ? a2 : b2;
		res = lds + 2 * MAXP;
		res[0].x = dot(u, imax); res[0].y = dot(w, imax); res[0].magic = 0x0;
		res[1].x = dot(u, imin); res[1].y = dot(w, imin); res[1].magic = 0x1;
		nr = 2;
	} else res = clip_areas(lds, na, sa, nb, sb, &nr);
	V3 amax = dot(A.ax, n) > 0 ? disk_extrema(cyl_top_center(A), A.ax, A.r, n) : disk_extrema(cyl_bot_center(A), A.ax, A.r, n);
	emit_manifold(res, nr, n, u, w, amax, mp.pen, pair_cid(ia, ib), false, P, slot, M, cap);
}

// detectSphereCylinder (CollisionDetect.cpp:1712-1792), lane 0.  The second
// disk centre repeats the top one (+h/2), as in the reference.
__device__ void sphere_cyl(const Bodies& b, const Shapes& sh, int is_, int ic, bool swapped, int slot, Manifolds& M, int cap) {
	V3 c = ldv(b.wpts + 3 * b.wpt_off[is_], 0);
	double r = sh.radius[b.shape[is_]];
	Cyl B = make_cyl_g(b, sh, ic);
	Check mp{-1, -1, V3{0, 0, 0}, __builtin_inf()};
	V3 dirs[4];
	V3 diff = B.c - c;
	dirs[0] = B.ax;
	dirs[1] = normalize(diff - B.ax * dot(B.ax, diff));
	V3 tdc = B.c + B.ax * 0.5 * B.h;
	V3 tcp = disk_extrema(tdc, B.ax, B.r, -normalize(diff));
	dirs[2] = normalize(tcp - c);
	dirs[3] = dirs[2];  // bottom disk centre evaluates to the top one
	for (int i = 0; i < 4; i++) {
		Check x = sat_check(sphere_extrema(c, r, dirs[i]), cyl_extrema(B, dirs[i]), dirs[i]);
		if (x.pen < 0) return;
		if (x.pen < mp.pen) mp = x;
	}
	if (!(mp.pen > 0)) return;
	V3 n = mp.n;
	V3 pos = c + n * r;
	uint64_t cid = pair_cid(is_, ic), bid = 0;
	if (swapped) {
		n = -n;
		cid = swap_halves(cid);
	}
	M.normal[slot] = n.x; M.normal[cap + slot] = n.y; M.normal[2 * cap + slot] = n.z;
	double* mpts = M.pts + (size_t)slot * 16;
	mpts[0] = pos.x; mpts[1] = pos.y; mpts[2] = pos.z; mpts[3] = mp.pen;
	M.magic[(size_t)slot * 8] = cid;
	M.magic[(size_t)slot * 8 + 1] = bid;
	M.maxpen[slot] = mp.pen;
	M.count[slot] = 1;
	write_mat(b, is_, ic, M, slot, cap);
}

// Removes non-finite points from a written manifold slot (lane 0); a
// non-finite normal drops the whole manifold.
__device__ void sanitize_slot(Manifolds& M, int slot, int cap) {
	int n = M.count[slot];
	if (n <= 0) return;
	double nx = M.normal[slot], ny = M.normal[cap + slot], nz = M.normal[2 * cap + slot];
	if (!isfinite(nx) || !isfinite(ny) || !isfinite(nz)) { M.count[slot] = 0; return; }
	double* p = M.pts + (size_t)slot * 16;
	uint64_t* g = M.magic + (size_t)slot * 8;
	int o = 0;
	for (int k = 0; k < n; k++) {
		if (!(isfinite(p[4 * k]) && isfinite(p[4 * k + 1]) && isfinite(p[4 * k + 2]) && isfinite(p[4 * k + 3]))) continue;
		if (o != k) {
			double* mt = M.mat + (size_t)slot * 12;
			for (int c = 0; c < 4; c++) p[4 * o + c] = p[4 * k + c];
			for (int c = 0; c < 3; c++) mt[3 * o + c] = mt[3 * k + c];
			g[2 * o] = g[2 * k];
			g[2 * o + 1] = g[2 * k + 1];
		}
		o++;
	}
	M.count[slot] = o;
}

// ------------------------------------------------------------- kernels
// Pairs are sorted into per-class lists first, so that each class runs in a
// kernel of its own shape: sphere-sphere and sphere-cylinder one lane per
// pair, the SAT classes one wave per pair.
enum { NPC_HH = 0, NPC_HS, NPC_SS, NPC_HC, NPC_SC, NPC_CC, NPC_MESH, NPC_N };

__global__ void k_np_zero(int* counts) {
	if (threadIdx.x < NPC_N) counts[threadIdx.x] = 0;
}

// Initialises every slot to "no manifold" and appends the pairs that need a
// test (PhysicsEngine.cpp:174: at least one dynamic body) to their class list;
// one atomic per class per wave.  List order does not matter: every slot's
// result depends on its pair only.
__device__ __forceinline__ bool sorted_has(const uint64_t* v, int n, uint64_t x) {
	int lo = 0, hi = n;
	while (lo < hi) {
		int mid = (lo + hi) >> 1;
		if (v[mid] < x) lo = mid + 1;
		else hi = mid;
	}
	return lo < n && v[lo] == x;
}

#define CLS_WAVES 16  // waves per classify block
__global__ __launch_bounds__(64 * CLS_WAVES) void k_np_classify(Bodies b, Shapes sh, const uint64_t* pairs, int npairs, Manifolds M, int* lists, int* counts,
                              const uint64_t* nocol, int n_nocol, const int* rank) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	int cls = -1;
	if (i < npairs) {
		M.count[i] = 0;
		M.maxpen[i] = -1;
		uint64_t key = pairs[i];
		int ia = (int)(key >> 32), ib = (int)(key & 0xFFFFFFFFu);
		// collisionAllowed (PhysicsEngine.cpp:570-572): jointed pairs never collide
		bool allowed = n_nocol == 0 || !sorted_has(nocol, n_nocol, key);
		// collision_possible (PhysicsEngine.cpp:174): a dynamic body, not both asleep
		bool both_asleep = b.asleep && b.asleep[ia] && b.asleep[ib];
		if (allowed && !(b.movement[ia] != 0 && b.movement[ib] != 0) && !both_asleep) {
			int ta = sh.type[b.shape[ia]], tb = sh.type[b.shape[ib]];
			int lo = ta < tb ? ta : tb, hi = ta < tb ? tb : ta;
			if (hi == SHAPE_MESH) cls = (lo != SHAPE_MESH && (!rank || rank[i] == 0)) ? NPC_MESH : -1;
			else if (lo == SHAPE_HULL) cls = hi == SHAPE_HULL ? NPC_HH : hi == SHAPE_SPHERE ? NPC_HS : NPC_HC;
			else if (lo == SHAPE_SPHERE) cls = hi == SHAPE_SPHERE ? NPC_SS : NPC_SC;
			else cls = NPC_CC;
		}
	}
	// list positions: wave ballots, a block-wide prefix per class in LDS, one
	// global atomic per class per block
	__shared__ int wcnt[NPC_N][CLS_WAVES];
	__shared__ int bbase[NPC_N];
	const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
	const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
	uint64_t mine = 0;
	for (int c = 0; c < NPC_N; c++) {
		uint64_t m = __ballot(cls == c);
		if (cls == c) mine = m;
		if (lane == 0) wcnt[c][wid] = __popcll(m);
	}
	__syncthreads();
	if (threadIdx.x < NPC_N) {
		int c = threadIdx.x, run = 0;
		for (int w = 0; w < CLS_WAVES; w++) {
			int v = wcnt[c][w];
			wcnt[c][w] = run;
			run += v;
		}
		bbase[c] = run ? atomicAdd(&counts[c], run) : 0;
	}
	__syncthreads();
	if (cls >= 0) lists[(size_t)cls * npairs + bbase[cls] + wcnt[cls][wid] + __popcll(mine & below)] = i;
}

#define CAND_K 4
// Wave-wide SAT with compaction: the cheap tests (pruning) run on 64 axes at
// a time and the survivors are appended to cand[] in axis order; whenever 64
// or more are waiting, whole batches of 64 are evaluated (the remainder moves
// to the front), and the tail is evaluated at the end, so the extrema queries
// keep every lane busy.  Each lane sees its axes in increasing order.  A
// separating axis anywhere ends the test (returns false); otherwise *bi is
// the first strict minimum (ties to the lower axis index), as in the
// reference's sequential loops.  cand[] needs 128 entries.
template <class Prune, class Eval>
__device__ bool wave_sat_compact(int naxes, int lane, int* cand, Prune prune, Eval eval, int* bi_out) {
	double bp = __builtin_inf();
	int bi = 0x7FFFFFFF;
	const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
	int nc = 0;
	for (int base = 0; base < naxes; base += 64) {
		int t = base + lane;
		bool v = t < naxes && prune(t);
		uint64_t m = __ballot(v);
		if (v) cand[nc + __popcll(m & below)] = t;
		nc += __popcll(m);
		const bool last = base + 64 >= naxes;
		if (nc < 64 && !(last && nc > 0)) continue;
		wave_fence();
		const int full = last ? nc : (nc & ~63);
		for (int c0 = 0; c0 < full; c0 += 64) {
			int idx = c0 + lane;
			int sep = 0;
			if (idx < full) {
				int a = cand[idx];
				Check x = eval(a);
				if (x.pen < 0) sep = 1;
				if (x.pen < bp) { bp = x.pen; bi = a; }
			}
			if (__ballot(sep) != 0) return false;
		}
		const int rem = nc - full;  // < 64
		int r = lane < rem ? cand[full + lane] : 0;
		wave_fence();
		if (lane < rem) cand[lane] = r;
		nc = rem;
		wave_fence();
	}
	for (int off = 32; off >= 1; off >>= 1) {
		double obp = __shfl_xor(bp, off);
		int obi = __shfl_xor(bi, off);
		if (obp < bp || (obp == bp && obi < bi)) { bp = obp; bi = obi; }
	}
	*bi_out = bi;
	return bp > 0 && bi != 0x7FFFFFFF;  // PhysicsEngine.cpp:213: keep only max_pen_depth > 0
}

// Copies a hull's world points and normals into the wave's LDS slice when they
// fit, so the extrema queries read LDS.
#define ST_PTS 32
#define ST_NRM 64
__device__ void stage_hull(Hull& h, double* spts, double* snrm, int lane) {
	if (h.np > ST_PTS || h.nf > ST_NRM) return;
	for (int k = lane; k < 3 * h.np; k += 64) spts[k] = h.wp[k];
	for (int k = lane; k < 3 * h.nf; k += 64) snrm[k] = h.wn[k];
	wave_fence();
	h.wp = spts;
	h.wn = snrm;
}

// Copies a hull's Gauss arc list (vertex index pairs) into LDS when it fits.
#define ST_ARC 128
__device__ void stage_arcs(Hull& h, int* sarc, int lane) {
	if (h.na > ST_ARC) return;
	for (int k = lane; k < 2 * h.na; k += 64) sarc[k] = h.arcs[k];
	wave_fence();
	h.arcs = sarc;
}

// Copies a cylinder's 14 world Gauss vertices into LDS.
__device__ void stage_cyl(Cyl& c, double* sgv, int lane) {
	if (lane < 42) sgv[lane] = c.gv[lane];
	wave_fence();
	c.gv = sgv;
}

// Arc-pair pruning of SAT_PolyPoly (CollisionDetect.cpp:750-770).
__device__ __forceinline__ bool hh_prune(const Hull& A, const Hull& B, int t) {
	if (t < A.nf) return !A.red[t];
	t -= A.nf;
	if (t < B.nf) return !B.red[t];
	t -= B.nf;
	int i1 = t / B.na, i2 = t - i1 * B.na;
	return arc_pair(ldv(A.wn, A.arcs[2 * i1]), ldv(A.wn, A.arcs[2 * i1 + 1]), -ldv(B.wn, B.arcs[2 * i2]), -ldv(B.wn, B.arcs[2 * i2 + 1]));
}

// A wave's LDS slice.  The axis search (candidate list, staged arcs) and the
// contact generation (clip polygons) never overlap in time, so they share
// storage; a wave_fence separates the last axis-search read from the first
// clip-polygon write.
struct SatLds {
	int cand[64 * CAND_K];
	int arca[2 * ST_ARC], arcb[2 * ST_ARC];
};
struct WaveLds {
	union {
		SatLds sat;
		CEP cep[3 * MAXP];
	};
	double pa[3 * ST_PTS], pb[3 * ST_PTS], na[3 * ST_NRM], nb[3 * ST_NRM];
};

// SAT_PolyPoly as a wave (CollisionDetect.cpp:717-788), same result as
// wave_sat_compact over eval_axis: the face axes of A then B, 64 at a time;
// then the arc pairs with each lane holding one or two of B's arcs in
// registers while A's arcs stream past uniformly (i1 outer, i2 inner, as the
// reference loops), surviving pairs compacted in axis order and evaluated 64
// at a time.  Any separating axis ends the test (returns false); otherwise
// *bi is the lowest-index axis of minimum depth.  Needs B.na <= 128.
__device__ bool wave_sat_hh(const Hull& A, const Hull& B, int lane, int* cand, int* bi_out) {
	double bp = __builtin_inf();
	int bi = 0x7FFFFFFF;
	const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
	auto consider = [&](int t) {
		bool v;
		Check x = eval_axis(A, B, t, &v);
		if (!v) return false;
		if (x.pen < bp || (x.pen == bp && t < bi)) {
			bp = x.pen;
			bi = t;
		}
		return x.pen < 0;
	};
	const int nfa = A.nf + B.nf;
	for (int base = 0; base < nfa; base += 64) {
		int t = base + lane;
		bool sep = false;
		if (t < nfa && !(t < A.nf ? A.red[t] : B.red[t - A.nf])) sep = consider(t);
		if (__ballot(sep) != 0) return false;
	}
	V3 b1[2], b2[2], bq[2];
	const int nh = (B.na + 63) >> 6;
#pragma unroll
	for (int h = 0; h < 2; h++) {
		int i2 = lane + 64 * h;
		if (h < nh && i2 < B.na) {
			b1[h] = -ldv(B.wn, B.arcs[2 * i2]);
			b2[h] = -ldv(B.wn, B.arcs[2 * i2 + 1]);
			bq[h] = cross(b1[h], b2[h]);
		}
	}
	int nc = 0;
	for (int i1 = 0; i1 < A.na; i1++) {
		V3 a1 = ldv(A.wn, A.arcs[2 * i1]), a2 = ldv(A.wn, A.arcs[2 * i1 + 1]);
		V3 aa = a1 + a2, ap = cross(a1, a2);
#pragma unroll
		for (int h = 0; h < 2; h++) {
			if (h >= nh) break;
			int i2 = lane + 64 * h;
			bool v = i2 < B.na && arc_test(a1, a2, aa, ap, b1[h], b2[h], bq[h]);
			uint64_t m = __ballot(v);
			if (v) cand[nc + __popcll(m & below)] = nfa + i1 * B.na + i2;
			nc += __popcll(m);
		}
		if (nc > 64 * CAND_K - 128 || (i1 == A.na - 1 && nc > 0)) {
			wave_fence();
			for (int c0 = 0; c0 < nc; c0 += 64) {
				int idx = c0 + lane;
				bool sep = idx < nc && consider(cand[idx]);
				if (__ballot(sep) != 0) return false;
			}
			nc = 0;
			wave_fence();
		}
	}
	for (int off = 32; off >= 1; off >>= 1) {
		double obp = __shfl_xor(bp, off);
		int obi = __shfl_xor(bi, off);
		if (obp < bp || (obp == bp && obi < bi)) { bp = obp; bi = obi; }
	}
	*bi_out = bi;
	return bp > 0 && bi != 0x7FFFFFFF;  // PhysicsEngine.cpp:213: keep only max_pen_depth > 0
}

// Waves per workgroup of the wave-per-pair kernels: single-wave groups let
// the concurrent pair classes pack the CUs most tightly (measured against 2
// and 4 on C3).
#define NP_WAVES 1
// Hull-hull: SAT_PolyPoly (CollisionDetect.cpp:717-835).
__global__ __launch_bounds__(64 * NP_WAVES) __attribute__((amdgpu_waves_per_eu(4, 4))) void k_np_hull_hull(Bodies b, Shapes sh, Params P, const uint64_t* pairs, const int* list,
                                                                const int* counts, Manifolds M, int cap) {
	__shared__ WaveLds lds[NP_WAVES];
	const int lane = threadIdx.x & 63, wib = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform
	WaveLds& L = lds[wib];
	const int cnt = counts[NPC_HH];
	for (int w = xcd_block() * NP_WAVES + wib; w < cnt; w += gridDim.x * NP_WAVES) {
		// one pair per wave: its indices are wave-uniform (scalar registers)
		int slot = __builtin_amdgcn_readfirstlane(list[w]);
		uint64_t key = pairs[slot];
		int ia = __builtin_amdgcn_readfirstlane((int)(key >> 32)), ib = __builtin_amdgcn_readfirstlane((int)(key & 0xFFFFFFFFu));
		Hull A = make_hull_g(b, sh, ia), B = make_hull_g(b, sh, ib);
		stage_hull(A, L.pa, L.na, lane);
		stage_hull(B, L.pb, L.nb, lane);
		stage_arcs(A, L.sat.arca, lane);
		stage_arcs(B, L.sat.arcb, lane);
		int naxes = A.nf + B.nf + A.na * B.na;
		int bi;
		bool hit = B.na <= 128 ? wave_sat_hh(A, B, lane, L.sat.cand, &bi)
		                       : wave_sat_compact(
		                             naxes, lane, L.sat.cand, [&](int t) { return hh_prune(A, B, t); },
		                             [&](int t) {
			                             bool v;
			                             return eval_axis(A, B, t, &v);
		                             },
		                             &bi);
		if (hit) {  // bi is wave-uniform
			bool valid;
			Check mp = eval_axis(A, B, bi, &valid);
			wave_fence();
			hull_contacts(A, B, mp, P, pair_cid(ia, ib), L.cep, slot, M, cap, lane);
			if (lane == 0) {
				write_mat(b, ia, ib, M, slot, cap);
				if (P.defect_fixes & 1) sanitize_slot(M, slot, cap);
			}
		}
		wave_fence();
	}
}

// Poly-cylinder pruning (degenerate edges, arc tests) for SAT_PolyCylinder.
__device__ __forceinline__ bool hc_prune(const Hull& A, const Cyl& B, int t) {
	if (t < A.nf) return !A.red[t];
	t -= A.nf;
	if (t == 0) return true;
	t -= 1;
	if (t < A.ne) {
		V3 ed = ldv(A.wp, A.edges[2 * t + 1]) - ldv(A.wp, A.edges[2 * t]);
		return !(mag(cross(ed, B.ax)) < 0.00000000001);
	}
	t -= A.ne;
	if (t < A.np) return true;
	t -= A.np;
	int i1 = t / 24, k = t - i1 * 24;
	return arc_pair(ldv(A.wn, A.arcs[2 * i1]), ldv(A.wn, A.arcs[2 * i1 + 1]), -cyl_gv(B, k & 1), -cyl_gv(B, k >> 1));
}

__global__ __launch_bounds__(64 * NP_WAVES) __attribute__((amdgpu_waves_per_eu(4, 4))) void k_np_hull_cyl(Bodies b, Shapes sh, Params P, const uint64_t* pairs, const int* list,
                                                               const int* counts, Manifolds M, int cap) {
	__shared__ WaveLds lds[NP_WAVES];
	const int lane = threadIdx.x & 63, wib = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform
	WaveLds& L = lds[wib];
	const int cnt = counts[NPC_HC];
	for (int w = xcd_block() * NP_WAVES + wib; w < cnt; w += gridDim.x * NP_WAVES) {
		// one pair per wave: its indices are wave-uniform (scalar registers)
		int slot = __builtin_amdgcn_readfirstlane(list[w]);
		uint64_t key = pairs[slot];
		int ia = __builtin_amdgcn_readfirstlane((int)(key >> 32)), ib = __builtin_amdgcn_readfirstlane((int)(key & 0xFFFFFFFFu));
		bool swapped = __builtin_amdgcn_readfirstlane(b.geo[3 * (size_t)ia].x) == SHAPE_CYLINDER;  // detectCollision's reversed form (CollisionDetect.cpp:122-130)
		int ip = swapped ? ib : ia, ic = swapped ? ia : ib;
		Hull A = make_hull_g(b, sh, ip);
		Cyl B = make_cyl_g(b, sh, ic);
		stage_hull(A, L.pa, L.na, lane);
		stage_arcs(A, L.sat.arca, lane);
		stage_cyl(B, L.nb, lane);
		int naxes = A.nf + 1 + A.ne + A.np + A.na * 24;
		int bi;
		bool hit = wave_sat_compact(
		    naxes, lane, L.sat.cand, [&](int t) { return hc_prune(A, B, t); },
		    [&](int t) {
			    bool v;
			    return eval_axis_pc(A, B, t, &v);
		    },
		    &bi);
		if (hit) {  // bi is wave-uniform
			bool valid;
			Check mp = eval_axis_pc(A, B, bi, &valid);
			wave_fence();
			poly_cyl_contacts(A, B, mp, P, ip, ic, swapped, L.cep, slot, M, cap, lane);
			if (lane == 0) {
				write_mat(b, ip, ic, M, slot, cap);
				if (P.defect_fixes & 1) sanitize_slot(M, slot, cap);
			}
		}
		wave_fence();
	}
}

__device__ __forceinline__ bool cc_prune(const Cyl& A, const Cyl& B, int t) {
	if (t < 2) return true;
	t -= 2;
	if (t < 32) {
		int g = t >> 3, i = t & 7;
		const double* poly = g == 0 ? A.top : g == 1 ? A.bot : g == 2 ? B.top : B.bot;
		V3 bax = g < 2 ? B.ax : A.ax;
		V3 ed = ldv(poly, (i + 1) % 8) - ldv(poly, i);
		return !(mag(cross(ed, bax)) < 0.00000000001);
	}
	t -= 32;
	if (t == 0) return mag2(cross(A.ax, B.ax)) > 0.00000001;
	t -= 1;
	int k1 = t / 24, k2 = t - k1 * 24;
	return arc_pair(cyl_gv(A, k1 & 1), cyl_gv(A, k1 >> 1), -cyl_gv(B, k2 & 1), -cyl_gv(B, k2 >> 1));
}

__global__ __launch_bounds__(64 * NP_WAVES) __attribute__((amdgpu_waves_per_eu(4, 4))) void k_np_cyl_cyl(Bodies b, Shapes sh, Params P, const uint64_t* pairs, const int* list,
                                                              const int* counts, Manifolds M, int cap) {
	__shared__ WaveLds lds[NP_WAVES];
	const int lane = threadIdx.x & 63, wib = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform
	WaveLds& L = lds[wib];
	const int cnt = counts[NPC_CC];
	for (int w = xcd_block() * NP_WAVES + wib; w < cnt; w += gridDim.x * NP_WAVES) {
		// one pair per wave: its indices are wave-uniform (scalar registers)
		int slot = __builtin_amdgcn_readfirstlane(list[w]);
		uint64_t key = pairs[slot];
		int ia = __builtin_amdgcn_readfirstlane((int)(key >> 32)), ib = __builtin_amdgcn_readfirstlane((int)(key & 0xFFFFFFFFu));
		Cyl A = make_cyl_g(b, sh, ia), B = make_cyl_g(b, sh, ib);
		stage_cyl(A, L.na, lane);
		stage_cyl(B, L.nb, lane);
		int bi;
		bool hit = wave_sat_compact(
		    35 + 24 * 24, lane, L.sat.cand, [&](int t) { return cc_prune(A, B, t); },
		    [&](int t) {
			    bool v;
			    return eval_axis_cc(A, B, t, &v);
		    },
		    &bi);
		// a hit leaves its axis for k_np_cyl_cyl_contacts (count -2 - axis)
		if (hit && lane == 0) M.count[slot] = -2 - bi;
		wave_fence();
	}
}

// Contact generation of the cylinder pairs k_np_cyl_cyl found touching, one
// wave per pair of the same list, lane 0 (kept out of the axis-search kernel
// so that kernel needs fewer registers and holds more waves).
__global__ __launch_bounds__(64 * NP_WAVES) void k_np_cyl_cyl_contacts(Bodies b, Shapes sh, Params P, const uint64_t* pairs, const int* list,
                                                                       const int* counts, Manifolds M, int cap) {
	__shared__ CEP cep[NP_WAVES][3 * MAXP];
	__shared__ double sgv[NP_WAVES][2][42];
	const int lane = threadIdx.x & 63, wib = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform
	const int cnt = counts[NPC_CC];
	for (int w = xcd_block() * NP_WAVES + wib; w < cnt; w += gridDim.x * NP_WAVES) {
		int slot = __builtin_amdgcn_readfirstlane(list[w]);
		const int code = __builtin_amdgcn_readfirstlane(M.count[slot]);
		if (code >= -1) continue;
		uint64_t key = pairs[slot];
		int ia = __builtin_amdgcn_readfirstlane((int)(key >> 32)), ib = __builtin_amdgcn_readfirstlane((int)(key & 0xFFFFFFFFu));
		Cyl A = make_cyl_g(b, sh, ia), B = make_cyl_g(b, sh, ib);
		stage_cyl(A, sgv[wib][0], lane);
		stage_cyl(B, sgv[wib][1], lane);
		if (lane == 0) {
			bool valid;
			Check mp = eval_axis_cc(A, B, -2 - code, &valid);
			cyl_cyl_contacts(A, B, mp, P, ia, ib, cep[wib], slot, M, cap);
			write_mat(b, ia, ib, M, slot, cap);
			if (P.defect_fixes & 1) sanitize_slot(M, slot, cap);
		}
		wave_fence();
	}
}

__global__ __launch_bounds__(64 * NP_WAVES) void k_np_hull_sphere(Bodies b, Shapes sh, Params P, const uint64_t* pairs, const int* list,
                                                                  const int* counts, Manifolds M, int cap) {
	__shared__ double sp[NP_WAVES][3 * ST_PTS], sn[NP_WAVES][3 * ST_NRM];
	const int lane = threadIdx.x & 63, wib = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform
	const int cnt = counts[NPC_HS];
	for (int w = xcd_block() * NP_WAVES + wib; w < cnt; w += gridDim.x * NP_WAVES) {
		// one pair per wave: its indices are wave-uniform (scalar registers)
		int slot = __builtin_amdgcn_readfirstlane(list[w]);
		uint64_t key = pairs[slot];
		int ia = __builtin_amdgcn_readfirstlane((int)(key >> 32)), ib = __builtin_amdgcn_readfirstlane((int)(key & 0xFFFFFFFFu));
		bool swapped = __builtin_amdgcn_readfirstlane(b.geo[3 * (size_t)ia].x) == SHAPE_SPHERE;
		const int ip = swapped ? ib : ia;
		Hull A = make_hull_g(b, sh, ip);
		stage_hull(A, sp[wib], sn[wib], lane);
		poly_sphere(b, sh, A, ip, swapped ? ia : ib, swapped, lane, slot, M, cap);
		wave_fence();
		if (lane == 0 && (P.defect_fixes & 1)) sanitize_slot(M, slot, cap);
	}
}

// Sphere-sphere and sphere-cylinder: one lane per pair.
__global__ void k_np_sphere_pairs(Bodies b, Shapes sh, Params P, const uint64_t* pairs, const int* list_ss, const int* list_sc,
                                  const int* counts, Manifolds M, int cap) {
	const int nss = counts[NPC_SS], nsc = counts[NPC_SC];
	const int stride = gridDim.x * blockDim.x;
	for (int w = blockIdx.x * blockDim.x + threadIdx.x; w < nss + nsc; w += stride) {
		int slot = w < nss ? list_ss[w] : list_sc[w - nss];
		uint64_t key = pairs[slot];
		int ia = (int)(key >> 32), ib = (int)(key & 0xFFFFFFFFu);
		if (w < nss) sphere_sphere(b, sh, ia, ib, slot, M, cap);
		else if (sh.type[b.shape[ia]] == SHAPE_SPHERE) sphere_cyl(b, sh, ia, ib, false, slot, M, cap);
		else sphere_cyl(b, sh, ib, ia, true, slot, M, cap);
		if (P.defect_fixes & 1) sanitize_slot(M, slot, cap);
	}
}

// ------------------------------------------------------------- static meshes
// A convex body against a static triangle mesh (SAT_{Poly,Sphere,Cylinder}Mesh,
// CollisionDetect.cpp:2647-2778, with the engine's merge and cull,
// PhysicsEngine.cpp:223-277): the triangles whose boxes meet the body's box,
// in the order of the mesh AABB tree's query; one wave per triangle runs the
// SAT (keeping both the overall and the Gauss-valid minimum) and lane 0 the
// contact generation; then per pair the near-parallel manifolds are merged
// into the first and each is culled to four points into the pair's slots.
struct Tri {
	const double* r;  // MT_REC record
	const int* ri;    // MT_IREC record
};
__device__ __forceinline__ Tri make_tri(const MeshDev& md, int t) {
	return Tri{md.tri + (size_t)t * MT_REC, md.tri_i + (size_t)t * MT_IREC};
}
__device__ __forceinline__ V3 tri_n(const Tri& t) { return ldv(t.r, 0); }
__device__ __forceinline__ V3 tri_v(const Tri& t, int k) { return ldv(t.r + 3, k); }
__device__ __forceinline__ V3 tri_out(const Tri& t, int k) { return ldv(t.r + 12, k); }
__device__ __forceinline__ V3 tri_g(const Tri& t, int k) { return ldv(t.r + 21, k); }

__device__ Ext tri_extrema(const Tri& t, V3 dir) {  // findTriangleExtrema (CollisionDetect.cpp:2004-2021)
	Ext e{-1, -1, __builtin_inf(), -__builtin_inf()};
	for (int i = 0; i < 3; i++) {
		double v = dot(tri_v(t, i), dir);
		if (v < e.min_val) { e.min_pid = i; e.min_val = v; }
		if (v > e.max_val) { e.max_pid = i; e.max_val = v; }
	}
	return e;
}

// normalDirectionValid (CollisionDetect.cpp:2024-2060).
__device__ bool tri_normal_valid(const Tri& t, V3 n) {
	const double EPS = 0.0001;
	int ng = t.ri[3];
	switch (t.ri[4]) {
	case 0:
	case 1:
		for (int i = 0; i < ng; i++) {
			V3 inner = cross(tri_g(t, i), tri_g(t, (i + 1) % ng));
			if (dot(n, inner) < -EPS) return false;
		}
		break;
	case 2: {
		V3 an = cross(tri_g(t, 0), tri_g(t, 1));
		if (fabs(dot(n, an)) > EPS) return false;
		if (dot(n, cross(an, tri_g(t, 0))) < -EPS) return false;
		if (dot(n, cross(tri_g(t, 1), an)) < -EPS) return false;
		break;
	}
	case 3:
		if (dot(n, tri_n(t)) < 1 - EPS) return false;
		break;
	}
	return true;
}

// findTriangleContactArea (CollisionDetect.cpp:282-329).
__device__ int tri_contact_area(const Tri& t, V3 n, V3 p, int pid, V3 u, V3 w, double cos_tol, double sin_tol, CEP* out, int* surface) {
	if (1 - dot(tri_n(t), n) <= cos_tol) {
		for (int i = 0; i < 3; i++) {
			V3 v = tri_v(t, i);
			out[i].x = dot(v, u);
			out[i].y = dot(v, w);
			out[i].src = i;
		}
		*surface = t.ri[5];
		return 3;
	}
	*surface = -1;
	for (int i = 0; i < 3; i++) {
		if (i != pid && (i + 1) % 3 != pid) continue;
		V3 p1 = tri_v(t, i), p2 = tri_v(t, (i + 1) % 3);
		if (fabs(dot(normalize(p2 - p1), n)) <= sin_tol) {
			out[0].x = dot(p1, u); out[0].y = dot(p1, w); out[0].src = t.ri[i];
			out[1].x = dot(p2, u); out[1].y = dot(p2, w); out[1].src = t.ri[i];
			return 2;
		}
	}
	out[0].x = dot(p, u);
	out[0].y = dot(p, w);
	out[0].src = pid;
	return 1;
}

// Arc k of a triangle's Gauss map: normal to edge e's outward direction, then
// on to the reversed normal (CollisionDetect.cpp:2097-2103), e = k / 2.
__device__ __forceinline__ void tri_arc(const Tri& t, int k, V3* g1, V3* g2) {
	int e = k >> 1;
	if ((k & 1) == 0) { *g1 = tri_n(t); *g2 = tri_out(t, e); }
	else { *g1 = tri_out(t, e); *g2 = -tri_n(t); }
}

// Candidate axis t of SAT_PolyTriangle (:2062-2160): the triangle normal, the
// poly's non-redundant face normals, then poly arc x triangle arc pairs.
__device__ Check eval_axis_pt(const Hull& A, const Tri& T, int t, bool* valid) {
	*valid = true;
	if (t == 0) {
		V3 n = tri_n(T);
		return sat_check(find_extrema(A, n), tri_extrema(T, n), n);
	}
	t -= 1;
	if (t < A.nf) {
		if (A.red[t]) { *valid = false; return Check{}; }
		V3 n = ldv(A.wn, t);
		return sat_check(recentered(A, t), tri_extrema(T, n), n);
	}
	t -= A.nf;
	int i1 = t / 6, k = t - 6 * i1;
	V3 g1, g2, n;
	tri_arc(T, k, &g1, &g2);
	if (!arc_axis(ldv(A.wn, A.arcs[2 * i1]), ldv(A.wn, A.arcs[2 * i1 + 1]), -g1, -g2, &n)) { *valid = false; return Check{}; }
	return sat_check(find_extrema(A, n), tri_extrema(T, n), n);
}
// Candidate axis t of SAT_SphereTriangle (:2196-2250): the triangle normal,
// vertex directions, edge normals; *feat = the axis' sphere feature id.
__device__ Check eval_axis_st(V3 c, double r, const Tri& T, int t, uint32_t* feat) {
	V3 n;
	if (t == 0) {
		n = tri_n(T);
		*feat = 0xFFFFFFFFu;
	} else if (t < 4) {
		n = normalize(tri_v(T, t - 1) - c);
		*feat = (uint32_t)(t - 1);
	} else {
		int e = t - 4;
		V3 p1 = tri_v(T, e), p2 = tri_v(T, (e + 1) % 3);
		V3 dir = normalize(p2 - p1);
		V3 sample = p1 - c;
		n = normalize(sample - dir * dot(dir, sample));
		*feat = (uint32_t)T.ri[e];
	}
	return sat_check(sphere_extrema(c, r, n), tri_extrema(T, n), n);
}
// Candidate axis t of SAT_CylinderTriangle (:2466-2596): the height axis, the
// triangle normal, triangle edge x axis, vertex-to-axis directions, then
// cylinder arc x triangle arc pairs.
__device__ Check eval_axis_ct(const Cyl& A, const Tri& T, int t, bool* valid) {
	*valid = true;
	V3 n;
	if (t == 0) n = A.ax;
	else if (t == 1) n = tri_n(T);
	else if (t < 5) {
		int e = t - 2;
		V3 dir = cross(tri_v(T, (e + 1) % 3) - tri_v(T, e), A.ax);
		if (mag(dir) < 0.00000000001) { *valid = false; return Check{}; }
		n = normalize(dir);
	} else if (t < 8) {
		V3 diff = tri_v(T, t - 5) - A.c;
		n = normalize(diff - A.ax * dot(A.ax, diff));
	} else {
		t -= 8;
		int k1 = t / 6, k = t - 6 * k1;
		V3 g1, g2;
		tri_arc(T, k, &g1, &g2);
		if (!arc_axis(cyl_gv(A, k1 & 1), cyl_gv(A, k1 >> 1), -g1, -g2, &n)) { *valid = false; return Check{}; }
	}
	return sat_check(cyl_extrema(A, n), tri_extrema(T, n), n);
}

// Wave-parallel triangle SAT: any separating axis ends the test (false);
// otherwise *bi is the first strict minimum over all axes and *vi the first
// strict minimum over the axes whose reversed normal is Gauss-valid for the
// triangle (0x7FFFFFFF if none).
template <class Eval>
__device__ bool wave_sat_tri(int naxes, int lane, const Tri& T, Eval eval, int* bi_out, double* bp_out, int* vi_out, double* vp_out) {
	double bp = __builtin_inf(), vp = __builtin_inf();
	int bi = 0x7FFFFFFF, vi = 0x7FFFFFFF;
	for (int base = 0; base < naxes; base += 64) {
		int t = base + lane;
		int sep = 0;
		if (t < naxes) {
			bool v;
			Check x = eval(t, &v);
			if (v) {
				if (x.pen < 0) sep = 1;
				if (x.pen < bp) { bp = x.pen; bi = t; }
				if (x.pen < vp && tri_normal_valid(T, -x.n)) { vp = x.pen; vi = t; }
			}
		}
		if (__ballot(sep) != 0) return false;
	}
	for (int off = 32; off >= 1; off >>= 1) {
		double obp = __shfl_xor(bp, off), ovp = __shfl_xor(vp, off);
		int obi = __shfl_xor(bi, off), ovi = __shfl_xor(vi, off);
		if (obp < bp || (obp == bp && obi < bi)) { bp = obp; bi = obi; }
		if (ovp < vp || (ovp == vp && ovi < vi)) { vp = ovp; vi = ovi; }
	}
	*bi_out = bi;
	*bp_out = bp;
	*vi_out = vi;
	*vp_out = vp;
	return bi != 0x7FFFFFFF;
}

__device__ __forceinline__ bool box_overlap(const double* a, const double* b6, int bs) {  // AABB::intersects (AABB.h:10-14)
	return !(a[3] < b6[0] || a[0] > b6[3 * bs] || a[4] < b6[bs] || a[1] > b6[4 * bs] || a[5] < b6[2 * bs] || a[2] > b6[5 * bs]);
}

// Slots of each pair: MESH_SLOTS for a convex body against a mesh that may
// collide (one dynamic body, not a jointed pair), else 1.
__global__ void k_mesh_expand(Bodies b, MeshDev md, const uint64_t* pairs, int npairs, const uint64_t* nocol, int n_nocol, int* cnt,
                              const int* off, uint64_t* out, int* rank) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= npairs) return;
	uint64_t key = pairs[i];
	int c = 1;
	if (!off) {
		int ia = (int)(key >> 32), ib = (int)(key & 0xFFFFFFFFu);
		bool ma = md.inst_of_body[ia] >= 0, mb = md.inst_of_body[ib] >= 0;
		bool allowed = n_nocol == 0 || !sorted_has(nocol, n_nocol, key);
		if (ma != mb && allowed && (b.movement[ia] == 0 || b.movement[ib] == 0)) c = MESH_SLOTS;
		cnt[i] = c;
		return;
	}
	c = cnt[i];
	for (int r = 0; r < c; r++) {
		out[off[i] + r] = key;
		rank[off[i] + r] = r;
	}
}

// Candidate triangles of each mesh pair (AABBTree::getCollisionCandidatesWith,
// AABB_Tree.h:226-247: depth first, right child first), counted (off null)
// or written at off[j].
__global__ void k_mesh_candidates(Bodies b, MeshDev md, const uint64_t* pairs, const int* list, const int* counts, int* cnt, const int* off,
                                  MeshItems it, int* overflow) {
	int j = blockIdx.x * blockDim.x + threadIdx.x;
	if (j >= counts[NPC_MESH]) return;
	int slot = list[j];
	uint64_t key = pairs[slot];
	int ia = (int)(key >> 32), ib = (int)(key & 0xFFFFFFFFu);
	int m = md.inst_of_body[ia] >= 0 ? ia : ib, c = m == ia ? ib : ia;
	const int N = b.n;
	int inst = md.inst_of_body[m];
	int stk[MESH_STACK];
	int sp = 0, k = 0;
	stk[sp++] = md.root[inst];
	while (sp > 0) {
		int n = stk[--sp];
		const double* nb = md.node_box + 6 * (size_t)n;
		bool hit = box_overlap(nb, b.aabb + c, N);
		if (md.node_lr[2 * n] < 0) {
			if (hit) {
				if (off) {
					it.slot[off[j] + k] = slot;
					it.tri[off[j] + k] = md.node_lr[2 * n + 1];
				}
				k++;
			}
		} else if (hit) {
			if (sp + 2 > MESH_STACK) {
				atomicAdd(overflow, 1);
				break;
			}
			stk[sp++] = md.node_lr[2 * n];
			stk[sp++] = md.node_lr[2 * n + 1];
		}
	}
	if (!off) cnt[j] = k;
}

// One wave per candidate triangle: the convex body's SAT against the triangle
// and, on lane 0, its contact points (convex body as the first shape).
__global__ __launch_bounds__(64 * NP_WAVES) void k_mesh_triangles(Bodies b, Shapes sh, MeshDev md, Params P, const uint64_t* pairs,
                                                                  MeshItems it, int nitems) {
	__shared__ WaveLds lds[NP_WAVES];
	const int lane = threadIdx.x & 63, wib = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform
	WaveLds& L = lds[wib];
	const int N = b.n;
	for (int w = xcd_block() * NP_WAVES + wib; w < nitems; w += gridDim.x * NP_WAVES) {
		// one (body, triangle) item per wave: its indices are wave-uniform
		int slot = __builtin_amdgcn_readfirstlane(it.slot[w]);
		uint64_t key = pairs[slot];
		int ia = __builtin_amdgcn_readfirstlane((int)(key >> 32)), ib = __builtin_amdgcn_readfirstlane((int)(key & 0xFFFFFFFFu));
		int c = md.inst_of_body[ia] >= 0 ? ib : ia;
		Tri T = make_tri(md, __builtin_amdgcn_readfirstlane(it.tri[w]));
		int ct = sh.type[b.shape[c]];
		// non_gauss_valid_normal_penalty (CollisionDetect.cpp:2673, :2707, :2770)
		double tl = T.r[33] - T.r[30], t2 = T.r[34] - T.r[31], t3 = T.r[35] - T.r[32];
		double cl = b.aabb[3 * N + c] - b.aabb[c], c2 = b.aabb[4 * N + c] - b.aabb[N + c], c3 = b.aabb[5 * N + c] - b.aabb[2 * N + c];
		double tlong = (tl < ((t2 < t3) ? t3 : t2)) ? ((t2 < t3) ? t3 : t2) : tl;
		double clong = (cl < ((c2 < c3) ? c3 : c2)) ? ((c2 < c3) ? c3 : c2) : cl;
		double small = (tlong < clong) ? tlong : clong;
		double penalty = (ct == SHAPE_HULL ? 0 : 0.15) * small;
		uint64_t ck = (uint64_t)(c + 1) * 16;
		uint64_t cid = (0x00000000FFFFFFFFull & ck) | (0xFFFFFFFF00000000ull & ((uint64_t)T.ri[5] << 32));
		int bi, vi;
		double bp, vp;
		bool hit;
		Hull A;
		Cyl C;
		V3 sc;
		double sr = 0;
		if (ct == SHAPE_HULL) {
			A = make_hull(b, sh, c);
			stage_hull(A, L.pa, L.na, lane);
			hit = wave_sat_tri(1 + A.nf + 6 * A.na, lane, T, [&](int t, bool* v) { return eval_axis_pt(A, T, t, v); }, &bi, &bp, &vi, &vp);
		} else if (ct == SHAPE_CYLINDER) {
			C = make_cyl(b, sh, c);
			hit = wave_sat_tri(8 + 24 * 6, lane, T, [&](int t, bool* v) { return eval_axis_ct(C, T, t, v); }, &bi, &bp, &vi, &vp);
		} else {
			sc = ldv(b.wpts + 3 * b.wpt_off[c], 0);
			sr = sh.radius[b.shape[c]];
			hit = wave_sat_tri(7, lane, T, [&](int t, bool* v) {
				uint32_t f;
				*v = true;
				return eval_axis_st(sc, sr, T, t, &f);
			}, &bi, &bp, &vi, &vp);
		}
		if (lane == 0) {
			int cnt = 0;
			bool gauss = vp < bp + penalty;
			double pen = gauss ? vp : bp;
			if (hit && pen > 0) {
				double* hp = it.head + 4 * (size_t)w;
				double* pp = it.pts + (size_t)w * MESH_PTS * 4;
				uint64_t* mg = it.magic + (size_t)w * MESH_PTS * 2;
				bool vv;
				if (ct == SHAPE_SPHERE) {
					uint32_t fr, fm;
					Check raw = eval_axis_st(sc, sr, T, bi, &fr);
					Check mp = gauss ? eval_axis_st(sc, sr, T, vi, &fm) : raw;
					V3 pos = sc + mp.n * sr;
					pp[0] = pos.x; pp[1] = pos.y; pp[2] = pos.z; pp[3] = mp.pen;
					mg[0] = cid;
					mg[1] = (uint64_t)fr;
					hp[0] = mp.n.x; hp[1] = mp.n.y; hp[2] = mp.n.z; hp[3] = mp.pen;
					cnt = 1;
				} else {
					Check raw = ct == SHAPE_HULL ? eval_axis_pt(A, T, bi, &vv) : eval_axis_ct(C, T, bi, &vv);
					Check mp = !gauss ? raw : (ct == SHAPE_HULL ? eval_axis_pt(A, T, vi, &vv) : eval_axis_ct(C, T, vi, &vv));
					V3 n = mp.n;
					V3 u, wv;
					perp_basis(n, &u, &wv);
					V3 bmax = tri_v(T, raw.b_pid);
					V3 amax;
					int na, sa, nb, sb, nr;
					CEP* res;
					nb = tri_contact_area(T, raw.n * -1.0, bmax, raw.b_pid, u, wv, P.cos_tol, P.sin_tol, L.cep + MAXP, &sb);
					if (ct == SHAPE_HULL) {
						amax = ldv(A.wp, raw.a_pid);
						na = contact_area(A, raw.n, amax, raw.a_pid, u, wv, P.cos_tol, P.sin_tol, L.cep, &sa);
						res = clip_areas(L.cep, na, sa, nb, sb, &nr);
					} else {
						amax = dot(C.ax, n) > 0 ? disk_extrema(cyl_top_center(C), C.ax, C.r, n) : disk_extrema(cyl_bot_center(C), C.ax, C.r, n);
						int oa;
						na = cyl_contact_area(C, raw.n, u, wv, P.cos_tol, P.sin_tol, L.cep, &sa, &oa);
						if (oa == AREA_EDGE) {
							res = L.cep;
							res[0].magic = 0x0;
							nr = 1;
						} else res = clip_areas(L.cep, na, sa, nb, sb, &nr);
					}
					double a_dot = dot(amax, n);
					V3 noff = n * a_dot;
					if (nr > MESH_PTS) nr = MESH_PTS;
					for (int q = 0; q < nr; q++) {
						V3 pos = u * res[q].x + wv * res[q].y + noff;
						pp[4 * q] = pos.x; pp[4 * q + 1] = pos.y; pp[4 * q + 2] = pos.z;
						pp[4 * q + 3] = dot(pos, n) - a_dot + mp.pen;
						mg[2 * q] = cid;
						mg[2 * q + 1] = res[q].magic;
					}
					hp[0] = n.x; hp[1] = n.y; hp[2] = n.z; hp[3] = mp.pen;
					cnt = nr;
				}
				// material of the points (CollisionDetect.cpp:2183-2187)
				double ra = b.mat[c], rt = T.r[36];
				it.mat[3 * (size_t)w] = (ra < rt) ? rt : ra;
				it.mat[3 * (size_t)w + 1] = (b.mat[N + c] + T.r[37]) / 2.0;
				it.mat[3 * (size_t)w + 2] = (b.mat[2 * N + c] + T.r[38]) / 2.0;
			}
			it.count[w] = cnt;
		}
		wave_fence();
	}
}

// Per mesh pair: manifolds in triangle order, flipped when the mesh is the
// pair's first body (CollisionDetect.cpp:171-195); each unmerged manifold
// absorbs the later ones whose normal is within the cos tolerance of its
// running normal (merge_manifold, :2780-2793: normal (n1 + n2) normalised,
// points appended, deepest depth), then is culled to four points
// (cull_manifold, :2795-2830) into the pair's next slot.
__global__ void k_mesh_merge(Params P, const uint64_t* pairs, MeshDev md, const int* list, int nlist, const int* off, MeshItems it, Manifolds M,
                             int cap, int* overflow) {
	int j = blockIdx.x * blockDim.x + threadIdx.x;
	if (j >= nlist) return;
	int slot0 = list[j];
	uint64_t key = pairs[slot0];
	bool flip = md.inst_of_body[(int)(key >> 32)] >= 0;
	int i0 = off[j], i1 = off[j + 1];
	auto nrm = [&](int i) {
		V3 n{it.head[4 * (size_t)i], it.head[4 * (size_t)i + 1], it.head[4 * (size_t)i + 2]};
		return flip ? -n : n;
	};
	for (int i = i0; i < i1; i++) {
		it.owner[i] = it.count[i] > 0 ? -1 : -2;
		int c = it.count[i];
		it.avail[i] = c >= 64 ? ~0ull : ((1ull << c) - 1);
	}
	int q = 0;
	for (int a = i0; a < i1; a++) {
		if (it.owner[a] != -1) continue;
		it.owner[a] = a;
		V3 n = nrm(a);
		double mp = it.head[4 * (size_t)a + 3];
		int total = it.count[a];
		for (int bb = a + 1; bb < i1; bb++) {
			if (it.owner[bb] != -1) continue;
			V3 nb = nrm(bb);
			if (1 - dot(n, nb) < P.cos_tol) {
				n = normalize(n + nb);
				double mb = it.head[4 * (size_t)bb + 3];
				mp = (mp < mb) ? mb : mp;
				it.owner[bb] = a;
				total += it.count[bb];
			}
		}
		if (q >= MESH_SLOTS) {
			atomicAdd(overflow, 1);
			continue;
		}
		int slot = slot0 + q++;
		M.normal[slot] = n.x;
		M.normal[cap + slot] = n.y;
		M.normal[2 * cap + slot] = n.z;
		double* mpts = M.pts + (size_t)slot * 16;
		uint64_t* mmag = M.magic + (size_t)slot * 8;
		double* mmat = M.mat + (size_t)slot * 12;
		auto put = [&](int k, int i, int p) {
			const double* src = it.pts + ((size_t)i * MESH_PTS + p) * 4;
			for (int d = 0; d < 4; d++) mpts[4 * k + d] = src[d];
			uint64_t c1 = it.magic[((size_t)i * MESH_PTS + p) * 2], c2 = it.magic[((size_t)i * MESH_PTS + p) * 2 + 1];
			mmag[2 * k] = flip ? swap_halves(c1) : c1;
			mmag[2 * k + 1] = flip ? swap_halves(c2) : c2;
			for (int d = 0; d < 3; d++) mmat[3 * k + d] = it.mat[3 * (size_t)i + d];
		};
		if (total <= 4) {
			int k = 0;
			for (int i = a; i < i1; i++) {
				if (it.owner[i] != a) continue;
				for (int p = 0; p < it.count[i]; p++) put(k++, i, p);
			}
			M.count[slot] = total;
			M.maxpen[slot] = mp;
		} else {
			V3 cu, cw;
			perp_basis(n, &cu, &cw);
			double maxpen = 0;
			for (int k = 0; k < 4; k++) {
				V3 t = cu * P.cull_c[k] + cw * P.cull_s[k];
				double bv = -__builtin_inf();
				int bi = -1, bpnt = -1;
				for (int i = a; i < i1; i++) {
					if (it.owner[i] != a) continue;
					for (int p = 0; p < it.count[i]; p++) {
						if (!((it.avail[i] >> p) & 1)) continue;
						const double* src = it.pts + ((size_t)i * MESH_PTS + p) * 4;
						double v = dot(V3{src[0], src[1], src[2]}, t);
						if (v > bv) { bv = v; bi = i; bpnt = p; }
					}
				}
				if (bi < 0) { bi = a; bpnt = 0; }
				put(k, bi, bpnt);
				double pp = it.pts[((size_t)bi * MESH_PTS + bpnt) * 4 + 3];
				maxpen = (maxpen < pp) ? pp : maxpen;
				it.avail[bi] &= ~(1ull << bpnt);
			}
			M.count[slot] = 4;
			M.maxpen[slot] = maxpen;
		}
		if (P.defect_fixes & 1) sanitize_slot(M, slot, cap);
	}
}

void launch_mesh_expand(const Bodies& b, const MeshDev& md, const uint64_t* pairs, int npairs, const uint64_t* nocol, int n_nocol,
                        int* cnt, const int* off, uint64_t* out, int* rank, hipStream_t s) {
	if (npairs > 0)
		hipLaunchKernelGGL(k_mesh_expand, dim3((npairs + 255) / 256), dim3(256), 0, s, b, md, pairs, npairs, nocol, n_nocol, cnt, off, out,
		                   rank);
}
void launch_mesh_candidates(const Bodies& b, const MeshDev& md, const uint64_t* pairs, const int* list, const int* counts, int nlist,
                            int* cnt, const int* off, MeshItems it, int* overflow, hipStream_t s) {
	if (nlist > 0)
		hipLaunchKernelGGL(k_mesh_candidates, dim3((nlist + 127) / 128), dim3(128), 0, s, b, md, pairs, list, counts, cnt, off, it, overflow);
}
void launch_mesh_triangles(const Bodies& b, const Shapes& sh, const MeshDev& md, const Params& p, const uint64_t* pairs, MeshItems it,
                           int nitems, hipStream_t s) {
	if (nitems <= 0) return;
	int blocks = (nitems + NP_WAVES - 1) / NP_WAVES;
	if (blocks > 32768 / NP_WAVES) blocks = 32768 / NP_WAVES;
	blocks = xcd_grid(blocks);
	hipLaunchKernelGGL(k_mesh_triangles, dim3(blocks), dim3(64 * NP_WAVES), 0, s, b, sh, md, p, pairs, it, nitems);
}
void launch_mesh_merge(const Params& p, const uint64_t* pairs, const MeshDev& md, const int* list, int nlist, const int* off, MeshItems it,
                       Manifolds m, int cap, int* overflow, hipStream_t s) {
	if (nlist > 0)
		hipLaunchKernelGGL(k_mesh_merge, dim3((nlist + 127) / 128), dim3(128), 0, s, p, pairs, md, list, nlist, off, it, m, cap, overflow);
}

void launch_narrowphase(const Bodies& b, const Shapes& sh, const Params& p, const uint64_t* pairs, int npairs, const Manifolds& m,
                        int cap, int* lists, int* counts, const uint64_t* nocol, int n_nocol, const int* rank, hipStream_t s, const Fork* fk) {
	if (npairs <= 0) return;
	hipLaunchKernelGGL(k_np_zero, dim3(1), dim3(64), 0, s, counts);
	if (b.nl) hipLaunchKernelGGL(k_body_geo, dim3((b.nl + 255) / 256), dim3(256), 0, s, b, sh);
	hipLaunchKernelGGL(k_np_classify, dim3((npairs + 64 * CLS_WAVES - 1) / (64 * CLS_WAVES)), dim3(64 * CLS_WAVES), 0, s, b, sh, pairs, npairs,
	                   m, lists, counts, nocol, n_nocol, rank);
	int wblocks = (npairs + NP_WAVES - 1) / NP_WAVES;
	if (wblocks > 16384 / NP_WAVES) wblocks = 16384 / NP_WAVES;
	wblocks = xcd_grid(wblocks);
	const int* L = lists;
	size_t n = (size_t)npairs;
	// the pair classes touch disjoint slots and run side by side on three
	// streams, each stream a long SAT class followed by a short one:
	// hull-cylinder then hull-sphere, hull-hull then the sphere pairs, and
	// cylinder-cylinder alone (it keeps fewer waves per SIMD, so it runs long
	// while it shares the chip)
	hipStream_t s0 = fk ? fk->side[0] : s, s1 = fk ? fk->side[1] : s;
	if (fk) {
		(void)(hipEventRecord(fk->fork, s));
		(void)(hipStreamWaitEvent(s0, fk->fork, 0));
		(void)(hipStreamWaitEvent(s1, fk->fork, 0));
	}
	hipLaunchKernelGGL(k_np_hull_cyl, dim3(wblocks), dim3(64 * NP_WAVES), 0, s, b, sh, p, pairs, L + NPC_HC * n, counts, m, cap);
	hipLaunchKernelGGL(k_np_hull_hull, dim3(wblocks), dim3(64 * NP_WAVES), 0, s0, b, sh, p, pairs, L + NPC_HH * n, counts, m, cap);
	hipLaunchKernelGGL(k_np_cyl_cyl, dim3(wblocks), dim3(64 * NP_WAVES), 0, s1, b, sh, p, pairs, L + NPC_CC * n, counts, m, cap);
	hipLaunchKernelGGL(k_np_cyl_cyl_contacts, dim3(wblocks), dim3(64 * NP_WAVES), 0, s1, b, sh, p, pairs, L + NPC_CC * n, counts, m, cap);
	hipLaunchKernelGGL(k_np_hull_sphere, dim3(wblocks), dim3(64 * NP_WAVES), 0, s, b, sh, p, pairs, L + NPC_HS * n, counts, m, cap);
	int tblocks = (npairs + 255) / 256;
	if (tblocks > 4096) tblocks = 4096;
	hipLaunchKernelGGL(k_np_sphere_pairs, dim3(tblocks), dim3(256), 0, s0, b, sh, p, pairs, L + NPC_SS * n, L + NPC_SC * n, counts, m, cap);
	if (fk) {
		(void)(hipEventRecord(fk->join[0], s0));
		(void)(hipEventRecord(fk->join[1], s1));
		(void)(hipStreamWaitEvent(s, fk->join[0], 0));
		(void)(hipStreamWaitEvent(s, fk->join[1], 0));
	}
}

}  // namespace cpx
