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
// The contact feature search, 2-D Sutherland-Hodgman clipping with feature-id
// (MagicID) bookkeeping, lift to 3-D and cull to four points (CollisionDetect.cpp
// :237-280, :392-620, :809-831, :2795-2830) are inherently serial and short;
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

struct Ext {
	int min_pid, max_pid;
	double min_val, max_val;
};

__device__ Ext find_extrema(const Hull& h, V3 axis) {  // CollisionDetect.cpp:627-644
	Ext e{-1, -1, __builtin_inf(), -__builtin_inf()};
	for (int i = 0; i < h.np; i++) {
		double v = dot(ldv(h.wp, i), axis);
		if (v < e.min_val) { e.min_pid = i; e.min_val = v; }
		if (v > e.max_val) { e.max_pid = i; e.max_val = v; }
	}
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

// Serial contact generation for a hull pair after SAT (lane 0).
__device__ void hull_contacts(const Hull& A, const Hull& B, const Check& mp, const Params& P, uint64_t cid, CEP* lds, int slot,
                              Manifolds& M, int cap) {
	V3 n = mp.n;
	V3 amax = ldv(A.wp, mp.a_pid), bmax = ldv(B.wp, mp.b_pid);
	V3 u, w;
	perp_basis(n, &u, &w);
	int sa, sb;
	int na = contact_area(A, n, amax, mp.a_pid, u, w, P.cos_tol, P.sin_tol, lds, &sa);
	int nb = contact_area(B, n * -1.0, bmax, mp.b_pid, u, w, P.cos_tol, P.sin_tol, lds + MAXP, &sb);
	int nr;
	CEP* res = clip_areas(lds, na, sa, nb, sb, &nr);
	emit_manifold(res, nr, n, u, w, amax, mp.pen, cid, false, P, slot, M, cap);
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
	M.mat[slot] = (ra < rb) ? rb : ra;
	M.mat[cap + slot] = (b.mat[b.n + ia] + b.mat[b.n + ib]) / 2.0;
	M.mat[2 * cap + slot] = (b.mat[2 * b.n + ia] + b.mat[2 * b.n + ib]) / 2.0;
}

// Wave-parallel SAT_PolySphere; poly = body ip, sphere = body is_.  When the
// sphere is the pair's first body the manifold is flipped as detectCollision
// does (CollisionDetect.cpp:91-99): normal negated, MagicID halves swapped.
__device__ void poly_sphere(const Bodies& b, const Shapes& sh, int ip, int is_, bool swapped, int lane, int slot, Manifolds& M, int cap) {
	Hull A = make_hull(b, sh, ip);
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
// SAT, lane 0; swapped when the cylinder is the pair's first body.
__device__ void poly_cyl_contacts(const Hull& A, const Cyl& B, const Check& mp, const Params& P, int ip, int ic, bool swapped, CEP* lds,
                                  int slot, Manifolds& M, int cap) {
	V3 n = mp.n;
	V3 amax = ldv(A.wp, mp.a_pid);
	V3 u, w;
	perp_basis(n, &u, &w);
	int sa, sb, ob;
	int na = contact_area(A, n, amax, mp.a_pid, u, w, P.cos_tol, P.sin_tol, lds, &sa);
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
	Cyl B = make_cyl(b, sh, ic);
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
			for (int c = 0; c < 4; c++) p[4 * o + c] = p[4 * k + c];
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
enum { NPC_HH = 0, NPC_HS, NPC_SS, NPC_HC, NPC_SC, NPC_CC, NPC_N };

__device__ __forceinline__ void wave_fence() {
	__builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
	__builtin_amdgcn_wave_barrier();
}

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

__global__ void k_np_classify(Bodies b, Shapes sh, const uint64_t* pairs, int npairs, Manifolds M, int* lists, int* counts,
                              const uint64_t* nocol, int n_nocol) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	int cls = -1;
	if (i < npairs) {
		M.count[i] = 0;
		M.maxpen[i] = -1;
		uint64_t key = pairs[i];
		int ia = (int)(key >> 32), ib = (int)(key & 0xFFFFFFFFu);
		// collisionAllowed (PhysicsEngine.cpp:570-572): jointed pairs never collide
		bool allowed = n_nocol == 0 || !sorted_has(nocol, n_nocol, key);
		if (allowed && !(b.movement[ia] != 0 && b.movement[ib] != 0)) {
			int ta = sh.type[b.shape[ia]], tb = sh.type[b.shape[ib]];
			int lo = ta < tb ? ta : tb, hi = ta < tb ? tb : ta;
			if (lo == SHAPE_HULL) cls = hi == SHAPE_HULL ? NPC_HH : hi == SHAPE_SPHERE ? NPC_HS : NPC_HC;
			else if (lo == SHAPE_SPHERE) cls = hi == SHAPE_SPHERE ? NPC_SS : NPC_SC;
			else cls = NPC_CC;
		}
	}
	const int lane = threadIdx.x & 63;
	const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
	for (int c = 0; c < NPC_N; c++) {
		uint64_t m = __ballot(cls == c);
		if (!m) continue;
		int leader = __ffsll((long long)m) - 1;
		int base = 0;
		if (lane == leader) base = atomicAdd(&counts[c], __popcll(m));
		base = __shfl(base, leader);
		if (cls == c) lists[(size_t)c * npairs + base + __popcll(m & below)] = i;
	}
}

#define CAND_K 4
// Wave-wide SAT with compaction: each round prunes 64*CAND_K candidate axes
// (cheap tests), packs the survivors into cand[] in axis order, then
// evaluates them 64 at a time, so the extrema queries keep every lane busy.
// A separating axis anywhere ends the test (returns false); otherwise *bi is
// the first strict minimum (ties to the lower axis index), as in the
// reference's sequential loops.
template <class Prune, class Eval>
__device__ bool wave_sat_compact(int naxes, int lane, int* cand, Prune prune, Eval eval, int* bi_out) {
	double bp = __builtin_inf();
	int bi = 0x7FFFFFFF;
	const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
	for (int base = 0; base < naxes; base += 64 * CAND_K) {
		int nc = 0;
		for (int j = 0; j < CAND_K; j++) {
			int t = base + j * 64 + lane;
			bool v = t < naxes && prune(t);
			uint64_t m = __ballot(v);
			if (v) cand[nc + __popcll(m & below)] = t;
			nc += __popcll(m);
		}
		wave_fence();
		for (int c0 = 0; c0 < nc; c0 += 64) {
			int idx = c0 + lane;
			int sep = 0;
			if (idx < nc) {
				int t = cand[idx];
				Check x = eval(t);
				if (x.pen < 0) sep = 1;
				if (x.pen < bp) { bp = x.pen; bi = t; }
			}
			if (__ballot(sep) != 0) return false;
		}
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

// Arc-pair pruning of SAT_PolyPoly (CollisionDetect.cpp:750-770).
__device__ __forceinline__ bool hh_prune(const Hull& A, const Hull& B, int t) {
	if (t < A.nf) return !A.red[t];
	t -= A.nf;
	if (t < B.nf) return !B.red[t];
	t -= B.nf;
	int i1 = t / B.na, i2 = t - i1 * B.na;
	V3 n;
	return arc_axis(ldv(A.wn, A.arcs[2 * i1]), ldv(A.wn, A.arcs[2 * i1 + 1]), -ldv(B.wn, B.arcs[2 * i2]), -ldv(B.wn, B.arcs[2 * i2 + 1]), &n);
}

struct WaveLds {
	CEP cep[4 * MAXP];
	int cand[64 * CAND_K];
	double pa[3 * ST_PTS], pb[3 * ST_PTS], na[3 * ST_NRM], nb[3 * ST_NRM];
};

#define NP_WAVES 4
// Hull-hull: SAT_PolyPoly (CollisionDetect.cpp:717-835).
__global__ __launch_bounds__(64 * NP_WAVES) void k_np_hull_hull(Bodies b, Shapes sh, Params P, const uint64_t* pairs, const int* list,
                                                                const int* counts, Manifolds M, int cap) {
	__shared__ WaveLds lds[NP_WAVES];
	const int lane = threadIdx.x & 63, wib = threadIdx.x >> 6;
	WaveLds& L = lds[wib];
	const int cnt = counts[NPC_HH];
	for (int w = blockIdx.x * NP_WAVES + wib; w < cnt; w += gridDim.x * NP_WAVES) {
		int slot = list[w];
		uint64_t key = pairs[slot];
		int ia = (int)(key >> 32), ib = (int)(key & 0xFFFFFFFFu);
		Hull A = make_hull(b, sh, ia), B = make_hull(b, sh, ib);
		stage_hull(A, L.pa, L.na, lane);
		stage_hull(B, L.pb, L.nb, lane);
		int naxes = A.nf + B.nf + A.na * B.na;
		int bi;
		bool hit = wave_sat_compact(
		    naxes, lane, L.cand, [&](int t) { return hh_prune(A, B, t); },
		    [&](int t) {
			    bool v;
			    return eval_axis(A, B, t, &v);
		    },
		    &bi);
		if (hit && lane == 0) {
			bool valid;
			Check mp = eval_axis(A, B, bi, &valid);
			hull_contacts(A, B, mp, P, pair_cid(ia, ib), L.cep, slot, M, cap);
			write_mat(b, ia, ib, M, slot, cap);
			if (P.defect_fixes & 1) sanitize_slot(M, slot, cap);
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
	V3 n;
	return arc_axis(ldv(A.wn, A.arcs[2 * i1]), ldv(A.wn, A.arcs[2 * i1 + 1]), -cyl_gv(B, k & 1), -cyl_gv(B, k >> 1), &n);
}

__global__ __launch_bounds__(64 * NP_WAVES) void k_np_hull_cyl(Bodies b, Shapes sh, Params P, const uint64_t* pairs, const int* list,
                                                               const int* counts, Manifolds M, int cap) {
	__shared__ WaveLds lds[NP_WAVES];
	const int lane = threadIdx.x & 63, wib = threadIdx.x >> 6;
	WaveLds& L = lds[wib];
	const int cnt = counts[NPC_HC];
	for (int w = blockIdx.x * NP_WAVES + wib; w < cnt; w += gridDim.x * NP_WAVES) {
		int slot = list[w];
		uint64_t key = pairs[slot];
		int ia = (int)(key >> 32), ib = (int)(key & 0xFFFFFFFFu);
		bool swapped = sh.type[b.shape[ia]] == SHAPE_CYLINDER;  // detectCollision's reversed form (CollisionDetect.cpp:122-130)
		int ip = swapped ? ib : ia, ic = swapped ? ia : ib;
		Hull A = make_hull(b, sh, ip);
		Cyl B = make_cyl(b, sh, ic);
		stage_hull(A, L.pa, L.na, lane);
		int naxes = A.nf + 1 + A.ne + A.np + A.na * 24;
		int bi;
		bool hit = wave_sat_compact(
		    naxes, lane, L.cand, [&](int t) { return hc_prune(A, B, t); },
		    [&](int t) {
			    bool v;
			    return eval_axis_pc(A, B, t, &v);
		    },
		    &bi);
		if (hit && lane == 0) {
			bool valid;
			Check mp = eval_axis_pc(A, B, bi, &valid);
			poly_cyl_contacts(A, B, mp, P, ip, ic, swapped, L.cep, slot, M, cap);
			write_mat(b, ip, ic, M, slot, cap);
			if (P.defect_fixes & 1) sanitize_slot(M, slot, cap);
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
	V3 n;
	return arc_axis(cyl_gv(A, k1 & 1), cyl_gv(A, k1 >> 1), -cyl_gv(B, k2 & 1), -cyl_gv(B, k2 >> 1), &n);
}

__global__ __launch_bounds__(64 * NP_WAVES) void k_np_cyl_cyl(Bodies b, Shapes sh, Params P, const uint64_t* pairs, const int* list,
                                                              const int* counts, Manifolds M, int cap) {
	__shared__ WaveLds lds[NP_WAVES];
	const int lane = threadIdx.x & 63, wib = threadIdx.x >> 6;
	WaveLds& L = lds[wib];
	const int cnt = counts[NPC_CC];
	for (int w = blockIdx.x * NP_WAVES + wib; w < cnt; w += gridDim.x * NP_WAVES) {
		int slot = list[w];
		uint64_t key = pairs[slot];
		int ia = (int)(key >> 32), ib = (int)(key & 0xFFFFFFFFu);
		Cyl A = make_cyl(b, sh, ia), B = make_cyl(b, sh, ib);
		int bi;
		bool hit = wave_sat_compact(
		    35 + 24 * 24, lane, L.cand, [&](int t) { return cc_prune(A, B, t); },
		    [&](int t) {
			    bool v;
			    return eval_axis_cc(A, B, t, &v);
		    },
		    &bi);
		if (hit && lane == 0) {
			bool valid;
			Check mp = eval_axis_cc(A, B, bi, &valid);
			cyl_cyl_contacts(A, B, mp, P, ia, ib, L.cep, slot, M, cap);
			write_mat(b, ia, ib, M, slot, cap);
			if (P.defect_fixes & 1) sanitize_slot(M, slot, cap);
		}
		wave_fence();
	}
}

__global__ __launch_bounds__(64 * NP_WAVES) void k_np_hull_sphere(Bodies b, Shapes sh, Params P, const uint64_t* pairs, const int* list,
                                                                  const int* counts, Manifolds M, int cap) {
	const int lane = threadIdx.x & 63, wib = threadIdx.x >> 6;
	const int cnt = counts[NPC_HS];
	for (int w = blockIdx.x * NP_WAVES + wib; w < cnt; w += gridDim.x * NP_WAVES) {
		int slot = list[w];
		uint64_t key = pairs[slot];
		int ia = (int)(key >> 32), ib = (int)(key & 0xFFFFFFFFu);
		bool swapped = sh.type[b.shape[ia]] == SHAPE_SPHERE;
		poly_sphere(b, sh, swapped ? ib : ia, swapped ? ia : ib, swapped, lane, slot, M, cap);
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

void launch_narrowphase(const Bodies& b, const Shapes& sh, const Params& p, const uint64_t* pairs, int npairs, const Manifolds& m,
                        int cap, int* lists, int* counts, const uint64_t* nocol, int n_nocol, hipStream_t s) {
	if (npairs <= 0) return;
	hipLaunchKernelGGL(k_np_zero, dim3(1), dim3(64), 0, s, counts);
	hipLaunchKernelGGL(k_np_classify, dim3((npairs + 255) / 256), dim3(256), 0, s, b, sh, pairs, npairs, m, lists, counts, nocol, n_nocol);
	int wblocks = (npairs + NP_WAVES - 1) / NP_WAVES;
	if (wblocks > 4096) wblocks = 4096;
	const int* L = lists;
	size_t n = (size_t)npairs;
	hipLaunchKernelGGL(k_np_hull_hull, dim3(wblocks), dim3(64 * NP_WAVES), 0, s, b, sh, p, pairs, L + NPC_HH * n, counts, m, cap);
	hipLaunchKernelGGL(k_np_hull_cyl, dim3(wblocks), dim3(64 * NP_WAVES), 0, s, b, sh, p, pairs, L + NPC_HC * n, counts, m, cap);
	hipLaunchKernelGGL(k_np_cyl_cyl, dim3(wblocks), dim3(64 * NP_WAVES), 0, s, b, sh, p, pairs, L + NPC_CC * n, counts, m, cap);
	hipLaunchKernelGGL(k_np_hull_sphere, dim3(wblocks), dim3(64 * NP_WAVES), 0, s, b, sh, p, pairs, L + NPC_HS * n, counts, m, cap);
	int tblocks = (npairs + 255) / 256;
	if (tblocks > 4096) tblocks = 4096;
	hipLaunchKernelGGL(k_np_sphere_pairs, dim3(tblocks), dim3(256), 0, s, b, sh, p, pairs, L + NPC_SS * n, L + NPC_SC * n, counts, m, cap);
}

}  // namespace cpx
