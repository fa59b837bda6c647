// Batched raycasts against the world geometry of the last step: the query
// PhysicsEngine::raycastFirstIntersection answers (PhysicsEngine.cpp:582-609),
// for many rays at once.  One wave per ray: the lanes take the ray's scene's
// bodies 64 at a time, test each body's box (AABB::rayIntersectsAABB,
// AABB.h:143-173) and, where it is met, the body's geometry
// (RigidBody::checkRayIntersection, RigidBody.cpp:185-204): polyhedra face by
// face in surface order (ConvexPrimitive.cpp:731-750), spheres and cylinders
// analytically (:113-202, :300-317), static meshes through their AABB tree
// (AABB_Tree.h:249-270, Geometry.cpp:867-895).  The wave keeps the least
// distance, the lowest body index on a tie (the reference takes the first
// candidate of its body tree's traversal).  FP64 in the reference's
// operation order.
#include "cp_internal.h"

namespace cpx {

namespace {
struct Hit {
	bool hit;
	V3 p, n;
	double t;
};

__device__ bool ray_hits_box(double mn0, double mn1, double mn2, double mx0, double mx1, double mx2, V3 o, V3 d) {
	const double inf = __builtin_inf();
	double t_min = -inf, t_max = inf;
	const double eps = 0.00000000001;
	const double oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z}, mn[3] = {mn0, mn1, mn2}, mx[3] = {mx0, mx1, mx2};
	for (int i = 0; i < 3; i++) {
		if (fabs(dd[i]) < eps) {
			if (oo[i] < mn[i] || oo[i] > mx[i]) return false;
		} else {
			double t0 = (mn[i] - oo[i]) / dd[i];
			double t1 = (mx[i] - oo[i]) / dd[i];
			if (t0 > t1) {
				double tmp = t0;
				t0 = t1;
				t1 = tmp;
			}
			t_min = (t0 < t_min) ? t_min : t0;  // std::max<double>(t0, t_min)
			t_max = (t_max < t1) ? t_max : t1;  // std::min<double>(t1, t_max)
			if (t_max < 0 || t_max < t_min) return false;
		}
	}
	return true;
}

// The smaller non-negative root of a t^2 + b t + c (ConvexPrimitive.cpp:137-150, :179-192).
__device__ bool near_root(double a, double b, double c, double* t) {
	double radical = b * b - 4 * a * c;
	if (radical < 0) return false;
	if (radical == 0) {
		*t = -0.5 * b / a;
		return !(*t < 0);
	}
	double t1 = (-b + sqrt(radical)) / (2 * a);
	double t2 = (-b - sqrt(radical)) / (2 * a);
	if (t1 < 0 && t2 < 0) return false;
	if (t1 < 0) *t = t2;
	else if (t2 < 0) *t = t1;
	else *t = (t2 < t1) ? t2 : t1;
	return true;
}

__device__ Hit ray_sphere(V3 o, V3 d, V3 c, double r) {
	Hit h{false, {0, 0, 0}, {0, 0, 0}, 0};
	V3 rel = o - c;
	double a = mag2(d), b = 2 * dot(rel, d), cc = mag2(rel) - r * r, t;
	if (!near_root(a, b, cc, &t)) return h;
	h.hit = true;
	h.p = o + t * d;
	h.n = normalize(h.p - c);
	h.t = t;
	return h;
}

__device__ Hit ray_cylinder(V3 o, V3 d, V3 c, V3 ax, double r, double hgt) {
	Hit h{false, {0, 0, 0}, {0, 0, 0}, 0};
	V3 u = d - ax * dot(ax, d);
	V3 w = o - c - ax * (dot(ax, o) - dot(ax, c));
	double a = u.x * u.x + u.y * u.y + u.z * u.z;
	double b = 2 * (u.x * w.x + u.y * w.y + u.z * w.z);
	double cc = w.x * w.x + w.y * w.y + w.z * w.z - r * r;
	double t;
	if (near_root(a, b, cc, &t)) {
		V3 hp = o + t * d;
		double x = dot(hp - c, ax);
		if (!(fabs(x) > hgt / 2.0)) {
			h.hit = true;
			h.p = hp;
			h.n = normalize((hp - c) - ax * x);
			h.t = t;
		}
	}
	// the cap the ray can enter (checkDiskIntersection, :300-306)
	double sgn = dot(ax, d) < 0 ? 1.0 : -1.0;
	V3 dc = c + (sgn * ax) * (hgt / 2.0);
	double td = dot(ax, dc - o) / dot(ax, d);
	V3 ip = o + td * d;
	if (!(td < 0 || mag2(ip - dc) > r * r) && (!h.hit || td < h.t)) {
		h.hit = true;
		h.p = ip;
		h.n = ax;
		h.t = td;
	}
	return h;
}

__device__ V3 ldw(const double* p, int k) { return V3{p[3 * k], p[3 * k + 1], p[3 * k + 2]}; }

__device__ Hit ray_hull(const Shapes& sh, int s, const double* W, V3 o, V3 d) {
	Hit h{false, {0, 0, 0}, {0, 0, 0}, 0};
	for (int f = sh.face_off[s]; f < sh.face_off[s + 1]; f++) {
		const int* fv = sh.fv + sh.fv_off[f];
		const int m = sh.fv_off[f + 1] - sh.fv_off[f];
		V3 p1 = ldw(W, fv[0]), p2 = ldw(W, fv[1]), p3 = ldw(W, fv[sh.f_ncalc[f]]);
		V3 n = normalize(cross(p2 - p1, p3 - p2)) * (double)sh.f_ndir[f];  // Surface::normal on world points
		const double eps = 0.0001;
		if (dot(d, n) > -eps || dot(o - p1, n) < 0) continue;
		double t = dot(n, p1 - o) / dot(n, d);
		V3 ip = o + t * d;
		bool inside = true;
		for (int i = 0; i < m && inside; i++) {
			V3 e1 = ldw(W, fv[i]), e2 = ldw(W, fv[i + 1 == m ? 0 : i + 1]);
			if (dot(cross(e2 - e1, n), ip - e1) > 0) inside = false;
		}
		if (!inside) continue;
		h.hit = true;
		h.p = ip;
		h.n = n;
		h.t = t;
		return h;
	}
	return h;
}

__device__ Hit ray_mesh(const MeshDev& md, int inst, V3 o, V3 d) {
	Hit best{false, {0, 0, 0}, {0, 0, 0}, 0};
	int st[MESH_STACK];
	int sp = 0;
	st[sp++] = md.root[inst];
	while (sp > 0) {
		int nd = st[--sp];
		const double* bx = md.node_box + 6 * (size_t)nd;
		int l = md.node_lr[2 * nd], r = md.node_lr[2 * nd + 1];
		if (!ray_hits_box(bx[0], bx[1], bx[2], bx[3], bx[4], bx[5], o, d)) continue;
		if (l >= 0) {
			if (sp + 2 > MESH_STACK) break;  // deeper than any tree the host accepts (depth <= 63)
			st[sp++] = l;
			st[sp++] = r;
			continue;
		}
		const double* T = md.tri + (size_t)r * MT_REC;  // leaf: (-1, triangle); its box is the triangle's
		V3 n = ldw(T, 0), v0 = ldw(T + 3, 0), v1 = ldw(T + 3, 1), v2 = ldw(T + 3, 2);
		if (fabs(dot(n, d)) < 0.0000000001) continue;
		double t = -dot(o - v1, n) / dot(d, n);
		if (t < 0 || (best.hit && best.t < t)) continue;
		V3 hp = o + t * d;
		if (dot(hp - v0, ldw(T + 12, 0)) > 0 || dot(hp - v1, ldw(T + 12, 1)) > 0 || dot(hp - v2, ldw(T + 12, 2)) > 0) continue;
		best.hit = true;
		best.p = hp;
		best.n = n;
		best.t = t;
	}
	return best;
}
}  // namespace

__global__ __launch_bounds__(64) void k_raycast(Bodies b, Shapes sh, MeshDev md, const int* scene_body_off, int n_rays, const double* org,
                                                const double* dir, const int* ray_scene, int* hit_body, double* hit7) {
	const int lane = threadIdx.x;
	for (int r = blockIdx.x; r < n_rays; r += gridDim.x) {
		const V3 o{org[3 * r], org[3 * r + 1], org[3 * r + 2]}, d{dir[3 * r], dir[3 * r + 1], dir[3 * r + 2]};
		const int s = ray_scene ? ray_scene[r] : -1;
		const int b0 = s < 0 ? 0 : scene_body_off[s], b1 = s < 0 ? b.nl : scene_body_off[s + 1];
		const int n = b.n;
		int bi = -1;
		Hit bh{false, {0, 0, 0}, {0, 0, 0}, 0};
		for (int i = b0 + lane; i < b1; i += 64) {
			if (!ray_hits_box(b.aabb[i], b.aabb[n + i], b.aabb[2 * n + i], b.aabb[3 * n + i], b.aabb[4 * n + i], b.aabb[5 * n + i], o, d))
				continue;
			const int s2 = b.shape[i];
			const double* W = b.wpts + 3 * (size_t)b.wpt_off[i];
			Hit h;
			switch (sh.type[s2]) {
			case SHAPE_HULL: h = ray_hull(sh, s2, W, o, d); break;
			case SHAPE_SPHERE: h = ray_sphere(o, d, ldw(W, 0), sh.radius[s2]); break;
			case SHAPE_CYLINDER: h = ray_cylinder(o, d, ldw(W, 0), ldw(b.wnorm + 3 * (size_t)b.wgv_off[i], 0), sh.radius[s2], sh.height[s2]); break;
			default: h = ray_mesh(md, md.inst_of_body[i], o, d); break;
			}
			if (h.hit && (bi < 0 || h.t < bh.t)) {  // a lane's bodies come in ascending index order
				bh = h;
				bi = i;
			}
		}
		// least distance over the wave, lowest body index on a tie
		double bt = bi < 0 ? __builtin_inf() : bh.t;
		int bb = bi < 0 ? 0x7FFFFFFF : bi;
		for (int off = 32; off >= 1; off >>= 1) {
			double ot = __shfl_xor(bt, off);
			int ob = __shfl_xor(bb, off);
			if (ot < bt || (ot == bt && ob < bb)) {
				bt = ot;
				bb = ob;
			}
		}
		const bool mine = bi >= 0 && bi == bb;
		if (lane == 0) hit_body[r] = bb == 0x7FFFFFFF ? -1 : bb;
		if (mine) {
			double* out = hit7 + 7 * (size_t)r;
			out[0] = bh.p.x; out[1] = bh.p.y; out[2] = bh.p.z;
			out[3] = bh.n.x; out[4] = bh.n.y; out[5] = bh.n.z;
			out[6] = bh.t;
		}
	}
}

void launch_raycast(const Bodies& b, const Shapes& sh, const MeshDev& md, const int* scene_body_off, int n_rays, const double* org,
                    const double* dir, const int* ray_scene, int* hit_body, double* hit7, hipStream_t s) {
	if (n_rays <= 0) return;
	int g = n_rays < 16384 ? n_rays : 16384;
	hipLaunchKernelGGL(k_raycast, dim3(g), dim3(64), 0, s, b, sh, md, scene_body_off, n_rays, org, dir, ray_scene, hit_body, hit7);
}

}  // namespace cpx
