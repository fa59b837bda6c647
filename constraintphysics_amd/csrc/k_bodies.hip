// Per-body kernels: gravity, substep integration (implicit gyroscopic
// rotation), inertia update and the end-of-step world-geometry refresh.
// One thread per body; all arithmetic FP64 in the reference's operation order
// (this file is compiled with -ffp-contract=off).
//
// Roofline: HBM-bound streaming over the body SoA.  Integrate reads and writes
// com/q/vel/angv/pvel/pang + world and reference inertia: 2*(3+4+3+3+3+3)*8 +
// 36*8 B = 592 B per body per substep (FP64).
#include "cp_internal.h"
#include "k_integrate.h"

namespace cpx {

// PhysicsEngine.cpp:84-86: v += g * dt once per step for awake dynamic bodies.
__global__ void k_gravity(Bodies b, Params p) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= b.nl || b.movement[i] != 0) return;
	V3 v = LD3(b.vel, i, b.n);
	v += V3{p.gx, p.gy, p.gz} * p.dt;
	ST3(b.vel, i, b.n, v);
}

// The step-start loop with sleeping (PhysicsEngine.cpp:65-87) for a dynamic
// body the loop finds awake: RigidBody::recordMovementState (RigidBody.cpp:
// 347-352) into the history ring, PhysicsEngine::bodySleepy (:1221-1254) and
// the sleep counters, then gravity.  FP64 in the reference's operation order.
__global__ void k_gravity_sleep(Bodies b, Params p, SleepDev S) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= b.nl || b.movement[i] != 0) return;
	if (S.proc ? !S.proc[i] : S.asleep[i] != 0) return;
	const int n = b.n;
	V3 v = LD3(b.vel, i, n), w = LD3(b.angv, i, n);
	int len = S.hlen[i], head = S.hhead[i];
	if (S.hcap > 0) {
		int slot;
		if (len < S.hcap) {
			slot = head + len;
			if (slot >= S.hcap) slot -= S.hcap;
			len++;
		} else {
			slot = head;  // the oldest entry leaves
			head = head + 1 == S.hcap ? 0 : head + 1;
		}
		double* h = S.hist + (size_t)slot * 6 * n;
		h[i] = v.x; h[n + i] = v.y; h[2 * n + i] = v.z;
		h[3 * n + i] = w.x; h[4 * n + i] = w.y; h[5 * n + i] = w.z;
		S.hlen[i] = len;
		S.hhead[i] = head;
	}
	bool sleepy = false;
	if (!S.disabled[i] && !(len == 0 || 2 * len < S.hist_need)) {
		V3 av{0, 0, 0}, aw{0, 0, 0};
		for (int j = 0; j < len; j++) {
			int slot = head + j;
			if (slot >= S.hcap) slot -= S.hcap;
			const double* h = S.hist + (size_t)slot * 6 * n;
			av += V3{h[i], h[n + i], h[2 * n + i]};
			aw += V3{h[3 * n + i], h[4 * n + i], h[5 * n + i]};
		}
		const double dn = (double)len;
		av = V3{av.x / dn, av.y / dn, av.z / dn};
		aw = V3{aw.x / dn, aw.y / dn, aw.z / dn};
		int last = head + len - 1;
		if (last >= S.hcap) last -= S.hcap;
		const double* hf = S.hist + (size_t)head * 6 * n;
		const double* hb = S.hist + (size_t)last * 6 * n;
		const double span = S.step_time * (len - 1);
		V3 acc{(hb[i] - hf[i]) / span, (hb[n + i] - hf[n + i]) / span, (hb[2 * n + i] - hf[2 * n + i]) / span};
		V3 aacc{(hb[3 * n + i] - hf[3 * n + i]) / span, (hb[4 * n + i] - hf[4 * n + i]) / span, (hb[5 * n + i] - hf[5 * n + i]) / span};
		sleepy = mag(av) <= S.vel_eps && mag(aw) <= S.ang_vel_eps && mag(acc) <= S.accel_eps && mag(aacc) <= S.ang_accel_eps;
	}
	if (sleepy) {
		S.counter[i] += S.step_time;
	} else {
		int t = S.ticks[i] + 1;
		if (t >= S.tick_threshold) {
			S.counter[i] = 0.0;
			t = 0;
		}
		S.ticks[i] = t;
	}
	v += V3{p.gx, p.gy, p.gz} * p.dt;
	ST3(b.vel, i, n, v);
}

// update_positions (PhysicsEngine.cpp:390-410) followed by
// RigidBody::updateInertiaTensor (RigidBody.cpp:270-279).  For the FP32
// solve it refreshes the body's FP32 inverse inertia too, so only a step's
// first substep needs k_iinv32 (bodies it skips keep their inertia).
__global__ __launch_bounds__(128) void k_integrate(Bodies b, Params p) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= b.nl || b.movement[i] != 0 || (b.asleep && b.asleep[i])) return;
	integrate_body(b, p, i, p.precision == CP_PREC_SPLIT ? b.iinv32 : nullptr);
}

// updateInertiaTensor for every body without touching the stored orientation
// (used when a state is uploaded without its world inertia).
__global__ __launch_bounds__(128) void k_update_inertia(Bodies b, int first, int count, const int* list) {
	int t = blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= count) return;
	int i = list ? list[t] : first + t;
	const int n = b.n;
	if (i >= n) return;
	Q q = q_normalize(ldq(b.q, i, n));
	M3 r = q_rot(q), rc = q_rot(q_conj(q));
	stm(b.I, i, n, r * ldm(b.Iref, i, n) * rc);
	stm(b.Iinv, i, n, r * ldm(b.Irefinv, i, n) * rc);
}

__device__ __forceinline__ double dmin(double a, double b) { return (b < a) ? b : a; }  // std::min<double>
__device__ __forceinline__ double dmax(double a, double b) { return (a < b) ? b : a; }  // std::max<double>

// RigidBody::updateGeometry (RigidBody.cpp:281-296): world hull points and
// Gauss normals (Polyhedron::recomputeFromReference, ConvexPrimitive.cpp:715-729)
// and the AABB (ConvexPrimitive.cpp:660-676).  only_moved: the end-of-step
// refresh (PhysicsEngine.cpp:424-438), which normalises the orientation first
// and skips bodies whose pose did not change in the last substep.
__global__ __launch_bounds__(128) void k_refresh_geometry(Bodies b, Shapes sh, int only_moved, int first, int count, const int* list) {
	int t = blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= count) return;
	int i = list ? list[t] : first + t;
	const int n = b.n;
	if (i >= n) return;
	Q q = ldq(b.q, i, n);
	V3 com = LD3(b.com, i, n);
	if (only_moved) {
		if (b.movement[i] != 0 || (b.asleep && b.asleep[i])) return;
		if (!(neq(com, LD3(b.prev_com, i, n)) || q_neq(q, ldq(b.prev_q, i, n)))) return;
		q = q_normalize(q);
		stq(b.q, i, n, q);
	}
	int s = b.shape[i];
	if (sh.type[s] == SHAPE_MESH) return;  // placed on the host with the mesh tree
	M3 R = q_rot(q);
	const double inf = __builtin_inf();
	double mn0 = inf, mn1 = inf, mn2 = inf, mx0 = -inf, mx1 = -inf, mx2 = -inf;
	if (sh.type[s] == SHAPE_HULL) {
		int p0 = sh.pt_off[s], np = sh.pt_off[s + 1] - p0;
		int w0 = b.wpt_off[i];
		for (int k = 0; k < np; k++) {
			const double* rp = sh.ref_pts + 3 * (p0 + k);
			V3 w = com + R * V3{rp[0], rp[1], rp[2]};
			double* wp = b.wpts + 3 * (w0 + k);
			wp[0] = w.x; wp[1] = w.y; wp[2] = w.z;
			mn0 = dmin(mn0, w.x); mn1 = dmin(mn1, w.y); mn2 = dmin(mn2, w.z);
			mx0 = dmax(mx0, w.x); mx1 = dmax(mx1, w.y); mx2 = dmax(mx2, w.z);
		}
		int f0 = sh.face_off[s], nf = sh.face_off[s + 1] - f0;
		int g0 = b.wgv_off[i];
		for (int k = 0; k < nf; k++) {
			const double* gn = sh.g_norm + 3 * (f0 + k);
			V3 w = R * V3{gn[0], gn[1], gn[2]};
			double* wp = b.wnorm + 3 * (g0 + k);
			wp[0] = w.x; wp[1] = w.y; wp[2] = w.z;
		}
	} else if (sh.type[s] == SHAPE_CYLINDER) {
		// Cylinder::recomputeFromReference and gen_AABB (ConvexPrimitive.cpp:272-298):
		// world points [centre, top 8-gon, bottom 8-gon] with bottom = top - h * axis,
		// normals [axis, 14 Gauss vertices]; AABB over top[0], bot[0], top[1], ...
		const double* rp = sh.ref_pts + 3 * sh.pt_off[s];
		const double* gn = sh.g_norm + 3 * sh.face_off[s];
		double* wp = b.wpts + 3 * b.wpt_off[i];
		double* wn = b.wnorm + 3 * b.wgv_off[i];
		V3 c = R * V3{rp[0], rp[1], rp[2]} + com;
		V3 ax = R * V3{gn[0], gn[1], gn[2]};
		double h = sh.height[s];
		wp[0] = c.x; wp[1] = c.y; wp[2] = c.z;
		wn[0] = ax.x; wn[1] = ax.y; wn[2] = ax.z;
		for (int k = 1; k < 15; k++) {
			V3 g = R * V3{gn[3 * k], gn[3 * k + 1], gn[3 * k + 2]};
			wn[3 * k] = g.x; wn[3 * k + 1] = g.y; wn[3 * k + 2] = g.z;
		}
		for (int k = 0; k < 8; k++) {
			const double* t = rp + 3 * (1 + k);
			V3 top = R * V3{t[0], t[1], t[2]} + com;
			V3 bot = top - ax * h;
			double* wt = wp + 3 * (1 + k);
			double* wb = wp + 3 * (9 + k);
			wt[0] = top.x; wt[1] = top.y; wt[2] = top.z;
			wb[0] = bot.x; wb[1] = bot.y; wb[2] = bot.z;
			mn0 = dmin(mn0, top.x); mn1 = dmin(mn1, top.y); mn2 = dmin(mn2, top.z);
			mx0 = dmax(mx0, top.x); mx1 = dmax(mx1, top.y); mx2 = dmax(mx2, top.z);
			mn0 = dmin(mn0, bot.x); mn1 = dmin(mn1, bot.y); mn2 = dmin(mn2, bot.z);
			mx0 = dmax(mx0, bot.x); mx1 = dmax(mx1, bot.y); mx2 = dmax(mx2, bot.z);
		}
	} else if (sh.type[s] == SHAPE_SPHERE) {
		// Sphere::recomputeFromReference and gen_AABB (ConvexPrimitive.cpp:433-443)
		const double* rp = sh.ref_pts + 3 * sh.pt_off[s];
		V3 c = com + R * V3{rp[0], rp[1], rp[2]};
		double* wp = b.wpts + 3 * b.wpt_off[i];
		wp[0] = c.x; wp[1] = c.y; wp[2] = c.z;
		double r = sh.radius[s];
		V3 mn = c - V3{r, r, r}, mx = c + V3{r, r, r};
		mn0 = mn.x; mn1 = mn.y; mn2 = mn.z;
		mx0 = mx.x; mx1 = mx.y; mx2 = mx.z;
	}
	b.aabb[i] = mn0; b.aabb[n + i] = mn1; b.aabb[2 * n + i] = mn2;
	b.aabb[3 * n + i] = mx0; b.aabb[4 * n + i] = mx1; b.aabb[5 * n + i] = mx2;
}

static inline int nblk(int n, int t) { return (n + t - 1) / t; }

void launch_gravity(const Bodies& b, const Params& p, hipStream_t s) {
	if (b.nl) hipLaunchKernelGGL(k_gravity, dim3(nblk(b.nl, 256)), dim3(256), 0, s, b, p);
}
void launch_gravity_sleep(const Bodies& b, const Params& p, const SleepDev& S, hipStream_t s) {
	if (b.nl) hipLaunchKernelGGL(k_gravity_sleep, dim3(nblk(b.nl, 256)), dim3(256), 0, s, b, p, S);
}
void launch_integrate(const Bodies& b, const Params& p, hipStream_t s) {
	if (b.nl) hipLaunchKernelGGL(k_integrate, dim3(nblk(b.nl, 128)), dim3(128), 0, s, b, p);
}
void launch_update_inertia(const Bodies& b, hipStream_t s) {
	if (b.n) hipLaunchKernelGGL(k_update_inertia, dim3(nblk(b.n, 128)), dim3(128), 0, s, b, 0, b.n, nullptr);
}
void launch_refresh_geometry(const Bodies& b, const Shapes& sh, int only_moved, hipStream_t s) {
	int cnt = only_moved ? b.nl : b.n;
	if (cnt) hipLaunchKernelGGL(k_refresh_geometry, dim3(nblk(cnt, 128)), dim3(128), 0, s, b, sh, only_moved, 0, cnt, nullptr);
}

// Non-finite detection (the reference only asserts, PhysicsEngine.cpp:
// 1121-1122): a dynamic body whose pose or velocity holds a NaN or an
// infinity sets its flag and the world's `any` flag.
__global__ void k_check_finite(Bodies b, int* body_flag, int* any) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= b.nl || b.movement[i] != 0) return;
	const int n = b.n;
	bool ok = true;
	for (int c = 0; c < 3; c++) ok = ok && isfinite(b.com[c * n + i]) && isfinite(b.vel[c * n + i]) && isfinite(b.angv[c * n + i]);
	for (int c = 0; c < 4; c++) ok = ok && isfinite(b.q[c * n + i]);
	if (ok) return;
	body_flag[i] = 1;
	any[0] = 1;
}
// Rows of many scenes gathered back to back: output row r belongs to scene k
// with off[k] <= r < off[k+1] and is body start[k] + r - off[k]; 13 doubles
// (com, q, vel, angv) per row; flag[k] set when one of its bodies was found
// non-finite at the end of the last step.
__global__ void k_gather_states(Bodies b, const int* start, const int* off, int nsc, const int* body_flag, double* out, int* flag) {
	const int r = blockIdx.x * blockDim.x + threadIdx.x;
	if (r >= off[nsc]) return;
	int lo = 0, hi = nsc - 1;
	while (lo < hi) {  // last scene with off <= r
		const int mid = (lo + hi + 1) >> 1;
		if (off[mid] <= r) lo = mid;
		else hi = mid - 1;
	}
	const int i = start[lo] + r - off[lo], n = b.n;
	double* o = out + 13 * (size_t)r;
	for (int c = 0; c < 3; c++) {
		o[c] = b.com[c * n + i];
		o[7 + c] = b.vel[c * n + i];
		o[10 + c] = b.angv[c * n + i];
	}
	for (int c = 0; c < 4; c++) o[3 + c] = b.q[c * n + i];
	if (body_flag[i]) flag[lo] = 1;
}
__global__ void k_fill_int(int* p, int n, int v) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) p[i] = v;
}
__global__ void k_zero_list(int* p, const int* list, int n) {
	int k = blockIdx.x * blockDim.x + threadIdx.x;
	if (k < n) p[list[k]] = 0;
}
void launch_zero_list(int* p, const int* list, int n, hipStream_t s) {
	if (n > 0) hipLaunchKernelGGL(k_zero_list, dim3(nblk(n, 256)), dim3(256), 0, s, p, list, n);
}
void launch_check_finite(const Bodies& b, int* body_flag, int* any, hipStream_t s) {
	if (b.nl) hipLaunchKernelGGL(k_check_finite, dim3(nblk(b.nl, 256)), dim3(256), 0, s, b, body_flag, any);
}
void launch_gather_states(const Bodies& b, const int* start, const int* off, int nsc, int rows, const int* body_flag, double* out, int* flag,
                          hipStream_t s) {
	if (rows > 0) hipLaunchKernelGGL(k_gather_states, dim3(nblk(rows, 256)), dim3(256), 0, s, b, start, off, nsc, body_flag, out, flag);
}
void launch_fill_int(int* p, int n, int v, hipStream_t s) {
	if (n > 0) hipLaunchKernelGGL(k_fill_int, dim3(nblk(n, 256)), dim3(256), 0, s, p, n, v);
}

// Writes body records (ints: index, shape, movement; doubles: mass, reference
// I[9], I^-1[9], restitution, mu_k, mu_s, com[3], q[4], vel[3], angv[3]) into
// the SoA arrays, then derives world inertia and geometry for those bodies.
__global__ void k_body_scatter(Bodies b, int count, const int* ri, const double* rd) {
	int t = blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= count) return;
	const int n = b.n;
	int i = ri[3 * t];
	if (i < 0 || i >= n) return;
	const double* r = rd + (size_t)BODY_REC * t;
	b.shape[i] = ri[3 * t + 1];
	b.movement[i] = ri[3 * t + 2];
	b.mass[i] = r[0];
	for (int c = 0; c < 9; c++) {
		b.Iref[c * n + i] = r[1 + c];
		b.Irefinv[c * n + i] = r[10 + c];
	}
	for (int c = 0; c < 3; c++) {
		b.mat[c * n + i] = r[19 + c];
		b.com[c * n + i] = r[22 + c];
		b.prev_com[c * n + i] = r[22 + c];
		b.vel[c * n + i] = r[29 + c];
		b.angv[c * n + i] = r[32 + c];
		b.pvel[c * n + i] = 0;
		b.pang[c * n + i] = 0;
	}
	for (int c = 0; c < 4; c++) {
		b.q[c * n + i] = r[25 + c];
		b.prev_q[c * n + i] = r[25 + c];
	}
}
void launch_body_scatter(const Bodies& b, const Shapes& sh, int count, const int* ri, const double* rd, const int* list, hipStream_t s) {
	if (count <= 0) return;
	hipLaunchKernelGGL(k_body_scatter, dim3(nblk(count, 128)), dim3(128), 0, s, b, count, ri, rd);
	hipLaunchKernelGGL(k_update_inertia, dim3(nblk(count, 128)), dim3(128), 0, s, b, 0, count, list);
	hipLaunchKernelGGL(k_refresh_geometry, dim3(nblk(count, 128)), dim3(128), 0, s, b, sh, 0, 0, count, list);
}

// Per-body scene index from the scene offsets (binary search).
__global__ void k_scene_of_body(int n, const int* sbo, int nscenes, int* out) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	int lo = 0, hi = nscenes;
	while (hi - lo > 1) {
		int m = (lo + hi) >> 1;
		if (sbo[m] <= i) lo = m;
		else hi = m;
	}
	out[i] = lo;
}
void launch_scene_of_body(int n, const int* sbo, int nscenes, int* out, hipStream_t s) {
	if (n) hipLaunchKernelGGL(k_scene_of_body, dim3(nblk(n, 256)), dim3(256), 0, s, n, sbo, nscenes, out);
}

}  // namespace cpx
