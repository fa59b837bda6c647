// Constraint rows and projected Gauss-Seidel.
//
// k_build_rows restates Contact::updateSolverConstraints
// (PersistentConstraint.cpp:201-215) -> ContactConstraint (ConstraintSolver.cpp:
// 240-272) and FrictionConstraint (:277-343), plus the per-solve target values
// (:265-268 and ConstraintSolver.h:78).  Matrix products keep NMat's
// accumulation order (NMat.h:50-60, :91-103) so FP64 rows match bit for bit.
//
// Two sweep schedules over the same rows (ConstraintSolver.cpp:80-213):
//  * k_pgs_ordered: the reference's sequential order (one thread per island,
//    rows in the island-search order supplied by the host) — the parity mode.
//  * k_pgs_colored: one workgroup per scene; edges (body pairs) are coloured
//    so that no two edges of a colour share a dynamic body, each thread runs
//    one edge's contacts in cache order, and colours are separated by a
//    workgroup barrier.  Gauss-Seidel per body, in a graph-coloured order.
//
// Roofline: HBM/L2-bound.  Per contact per sweep the algorithmic traffic is the
// row (~350 B FP64: J, M^-1 J^T, effective masses, targets, impulses) plus the
// gather/scatter of two bodies' velocity deltas.
#include "cp_internal.h"

namespace cpx {

#define MAXC CP_MAXC

__device__ __forceinline__ V3 ld3(const double* a, int i, int n) { return V3{a[i], a[n + i], a[2 * n + i]}; }

__global__ void k_build_rows(Bodies b, Contacts C, int nc, Params P) {
	int k = blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= nc || !C.live[k]) return;
	const int N = b.n, K = C.cap;
	int i1 = (int)(C.key[k] >> 32), i2 = (int)(C.key[k] & 0xFFFFFFFFu);
	Q q1{b.q[i1], b.q[N + i1], b.q[2 * N + i1], b.q[3 * N + i1]};
	Q q2{b.q[i2], b.q[N + i2], b.q[2 * N + i2], b.q[3 * N + i2]};
	V3 c1 = ld3(b.com, i1, N), c2 = ld3(b.com, i2, N);
	V3 lp1 = ld3(C.lp1, k, K), lp2 = ld3(C.lp2, k, K), ln1 = ld3(C.ln1, k, K);
	V3 p1 = q_rotate(q1, lp1) + c1;
	V3 p2 = q_rotate(q2, lp2) + c2;
	V3 n = q_rotate(q1, ln1);
	double adj = dot(p2 - p1, n) + C.pen[k];
	double cimp = C.imp_n[k] * 1.0;
	double f0 = C.imp_t[k] * 1.0, f1 = C.imp_t[K + k] * 1.0;
	double a1 = P.hardness * P.sub_dt, a2 = 1.0 / P.sub_dt;
	double pcc = (a2 < a1) ? a2 : a1;  // posCorrectCoeff, PersistentConstraint.cpp:199
	bool d1 = b.movement[i1] == 0, d2 = b.movement[i2] == 0;
	double Ma[6][6], Mb[6][6];
	{
		double ima = d1 ? 1.0 / b.mass[i1] : 0, imb = d2 ? 1.0 / b.mass[i2] : 0;
		for (int r = 0; r < 6; r++) for (int c = 0; c < 6; c++) { Ma[r][c] = 0; Mb[r][c] = 0; }
		for (int r = 0; r < 3; r++)
			for (int c = 0; c < 3; c++) {
				Ma[r][c] = (r == c ? 1.0 : 0.0) * ima;
				Mb[r][c] = (r == c ? 1.0 : 0.0) * imb;
				Ma[3 + r][3 + c] = d1 ? b.Iinv[(3 * r + c) * N + i1] : 0.0;
				Mb[3 + r][3 + c] = d2 ? b.Iinv[(3 * r + c) * N + i2] : 0.0;
			}
	}
	V3 rA = p1 - c1, rB = p1 - c2;
	// crossMat (NMat.cpp:62-70)
	double sa[3][3] = {{0, -rA.z, rA.y}, {rA.z, 0, -rA.x}, {-rA.y, rA.x, 0}};
	double sb[3][3] = {{0, -rB.z, rB.y}, {rB.z, 0, -rB.x}, {-rB.y, rB.x, 0}};
	double va[6], vb[6];
	{
		V3 v = d1 ? ld3(b.vel, i1, N) : V3{0, 0, 0}, w = d1 ? ld3(b.angv, i1, N) : V3{0, 0, 0};
		va[0] = v.x; va[1] = v.y; va[2] = v.z; va[3] = w.x; va[4] = w.y; va[5] = w.z;
		// getVel()/getAngVel() are zero only for FIXED bodies (RigidBody.cpp:174-181)
		if (b.movement[i1] != 1 && !d1) {
			V3 v2 = ld3(b.vel, i1, N), w2 = ld3(b.angv, i1, N);
			va[0] = v2.x; va[1] = v2.y; va[2] = v2.z; va[3] = w2.x; va[4] = w2.y; va[5] = w2.z;
		}
	}
	{
		V3 v = d2 ? ld3(b.vel, i2, N) : V3{0, 0, 0}, w = d2 ? ld3(b.angv, i2, N) : V3{0, 0, 0};
		vb[0] = v.x; vb[1] = v.y; vb[2] = v.z; vb[3] = w.x; vb[4] = w.y; vb[5] = w.z;
		if (b.movement[i2] != 1 && !d2) {
			V3 v2 = ld3(b.vel, i2, N), w2 = ld3(b.angv, i2, N);
			vb[0] = v2.x; vb[1] = v2.y; vb[2] = v2.z; vb[3] = w2.x; vb[4] = w2.y; vb[5] = w2.z;
		}
	}
	// ---- ContactConstraint
	{
		double nd[3] = {n.x, n.y, n.z}, ndn[3] = {-n.x, -n.y, -n.z};
		double ja[6], jb[6];
		for (int j = 0; j < 3; j++) {
			double s1 = 0, s2 = 0;
			for (int q = 0; q < 3; q++) { s1 += nd[q] * sa[q][j]; s2 += ndn[q] * sb[q][j]; }
			ja[j] = ndn[j]; ja[3 + j] = s1;
			jb[j] = nd[j]; jb[3 + j] = s2;
		}
		double ita[6], itb[6];
		for (int r = 0; r < 6; r++) {
			double s1 = 0, s2 = 0;
			for (int q = 0; q < 6; q++) { s1 += Ma[r][q] * ja[q]; s2 += Mb[r][q] * jb[q]; }
			ita[r] = s1; itb[r] = s2;
		}
		double t1 = 0, t2 = 0;
		for (int q = 0; q < 6; q++) { t1 += ja[q] * ita[q]; t2 += jb[q] * itb[q]; }
		double itv = t1 + t2;
		itv *= (1 + P.cfm);
		for (int r = 0; r < 6; r++) {
			C.ja[r * K + k] = ja[r]; C.jb[r * K + k] = jb[r];
			C.ita[r * K + k] = ita[r]; C.itb[r * K + k] = itb[r];
		}
		C.inv_n[k] = 1.0 / itv;
		double sp = adj - 0.05;
		sp = (sp < 0.0) ? 0.0 : sp;  // std::max<double>(0.0, x)
		C.ptgt_n[k] = sp * pcc;
		C.pimp_n[k] = 0;
		C.imp_n[k] = cimp;
		double x = 0, y = 0;
		for (int q = 0; q < 6; q++) { x += va[q] * ja[q]; y += vb[q] * jb[q]; }
		double cur = x + y;
		C.tgt_n[k] = (cur < -P.cutoff_vel) ? -(1 + C.bounce[k]) * cur : -cur;
	}
	// ---- FrictionConstraint
	{
		V3 su = ld3(C.fu, k, K), sw = ld3(C.fw, k, K);
		V3 u, w;
		perp_basis(n, &u, &w);
		double i0 = dot(u, su) * f0 + dot(u, sw) * f1;
		double i1v = dot(w, su) * f0 + dot(w, sw) * f1;
		double ud[3] = {u.x, u.y, u.z}, wd[3] = {w.x, w.y, w.z};
		double nu[3] = {-u.x, -u.y, -u.z}, nw[3] = {-w.x, -w.y, -w.z};
		double ja[2][6], jb[2][6];
		for (int j = 0; j < 3; j++) {
			double au = 0, bu = 0, aw = 0, bw = 0;
			for (int q = 0; q < 3; q++) {
				au += nu[q] * sa[q][j];
				bu += ud[q] * sb[q][j];
				aw += nw[q] * sa[q][j];
				bw += wd[q] * sb[q][j];
			}
			ja[0][j] = ud[j]; ja[0][3 + j] = au;
			jb[0][j] = nu[j]; jb[0][3 + j] = bu;
			ja[1][j] = wd[j]; ja[1][3 + j] = aw;
			jb[1][j] = nw[j]; jb[1][3 + j] = bw;
		}
		double ita[6][2], itb[6][2];
		for (int r = 0; r < 6; r++)
			for (int c = 0; c < 2; c++) {
				double s1 = 0, s2 = 0;
				for (int q = 0; q < 6; q++) { s1 += Ma[r][q] * ja[c][q]; s2 += Mb[r][q] * jb[c][q]; }
				ita[r][c] = s1; itb[r][c] = s2;
			}
		double itv[2][2];
		for (int r = 0; r < 2; r++)
			for (int c = 0; c < 2; c++) {
				double s1 = 0, s2 = 0;
				for (int q = 0; q < 6; q++) { s1 += ja[r][q] * ita[q][c]; s2 += jb[r][q] * itb[q][c]; }
				itv[r][c] = s1 + s2;
			}
		itv[0][0] *= (1 + P.cfm);
		itv[1][1] *= (1 + P.cfm);
		double a = itv[0][0], bq = itv[0][1], cq = itv[1][0], d = itv[1][1];
		double dt = a * d - bq * cq;
		double di = 1.0 / dt;
		C.finv[k] = d * di; C.finv[K + k] = -bq * di; C.finv[2 * K + k] = -cq * di; C.finv[3 * K + k] = a * di;
		for (int r = 0; r < 2; r++)
			for (int c = 0; c < 6; c++) { C.fja[(r * 6 + c) * K + k] = ja[r][c]; C.fjb[(r * 6 + c) * K + k] = jb[r][c]; }
		for (int r = 0; r < 6; r++)
			for (int c = 0; c < 2; c++) { C.fita[(r * 2 + c) * K + k] = ita[r][c]; C.fitb[(r * 2 + c) * K + k] = itb[r][c]; }
		C.imp_t[k] = i0;
		C.imp_t[K + k] = i1v;
		C.fu[k] = u.x; C.fu[K + k] = u.y; C.fu[2 * K + k] = u.z;
		C.fw[k] = w.x; C.fw[K + k] = w.y; C.fw[2 * K + k] = w.z;
		C.sready[k] = 0;
		for (int r = 0; r < 2; r++) {
			double x = 0, y = 0;
			for (int q = 0; q < 6; q++) { x += va[q] * ja[r][q]; y += vb[q] * jb[r][q]; }
			C.ftgt[r * K + k] = -(x + y);
		}
	}
}

// ------------------------------------------------------------- PGS row updates
struct DvRef {
	double* base;
	int i, n;
	bool dyn;
	__device__ double get(int c) const { return base[c * n + i]; }
	__device__ void add(int c, double v) const {
		if (dyn) base[c * n + i] += v;
	}
};

__device__ __forceinline__ void warm_start(const Contacts& C, int k, const DvRef& A, const DvRef& B) {
	const int K = C.cap;
	double imp = C.imp_n[k];
	if (!is_zero1(imp)) {
		for (int r = 0; r < 6; r++) {
			double da = 0, db = 0;
			da += imp * C.ita[r * K + k];
			db += imp * C.itb[r * K + k];
			A.add(r, da);
			B.add(r, db);
		}
	}
	double t0 = C.imp_t[k], t1 = C.imp_t[K + k];
	if (!is_zero2(t0, t1)) {
		for (int r = 0; r < 6; r++) {
			double da = 0, db = 0;
			da += t0 * C.fita[(r * 2) * K + k];
			da += t1 * C.fita[(r * 2 + 1) * K + k];
			db += t0 * C.fitb[(r * 2) * K + k];
			db += t1 * C.fitb[(r * 2 + 1) * K + k];
			A.add(r, da);
			B.add(r, db);
		}
	}
}

// PGSConstraintStep<1> for the contact row (ConstraintSolver.cpp:12-40, :270-272).
__device__ __forceinline__ void normal_step(Contacts& C, int k, const DvRef& A, const DvRef& B, bool pseudo) {
	const int K = C.cap;
	int off = pseudo ? 6 : 0;
	double x = 0, y = 0;
	for (int q = 0; q < 6; q++) {
		x += A.get(off + q) * C.ja[q * K + k];
		y += B.get(off + q) * C.jb[q * K + k];
	}
	double cur = x + y;
	double tgt = pseudo ? C.ptgt_n[k] : C.tgt_n[k];
	double acc = pseudo ? C.pimp_n[k] : C.imp_n[k];
	double delta = tgt - cur;
	double m = 0;
	m += delta * C.inv_n[k];
	double nw = acc + m;
	double pr = (nw < 0) ? 0 : nw;  // std::max<double>(nw, 0)
	double diff = pr - acc;
	if (!is_zero1(diff)) {
		for (int r = 0; r < 6; r++) {
			double da = 0, db = 0;
			da += diff * C.ita[r * K + k];
			db += diff * C.itb[r * K + k];
			A.add(off + r, da);
			B.add(off + r, db);
		}
		if (pseudo) C.pimp_n[k] = acc + diff;
		else C.imp_n[k] = acc + diff;
	}
}

// PGSConstraintStep<2> with FrictionConstraint::projectValidImpulse (ConstraintSolver.cpp:331-343).
__device__ __forceinline__ void friction_step(Contacts& C, int k, const DvRef& A, const DvRef& B) {
	const int K = C.cap;
	double cur[2];
	for (int r = 0; r < 2; r++) {
		double x = 0, y = 0;
		for (int q = 0; q < 6; q++) {
			x += A.get(q) * C.fja[(r * 6 + q) * K + k];
			y += B.get(q) * C.fjb[(r * 6 + q) * K + k];
		}
		cur[r] = x + y;
	}
	double d0 = C.ftgt[k] - cur[0], d1 = C.ftgt[K + k] - cur[1];
	double m0 = 0, m1 = 0;
	m0 += d0 * C.finv[k];
	m0 += d1 * C.finv[K + k];
	m1 += d0 * C.finv[2 * K + k];
	m1 += d1 * C.finv[3 * K + k];
	double i0 = C.imp_t[k], i1 = C.imp_t[K + k];
	double n0 = i0 + m0, n1 = i1 + m1;
	double ni = C.imp_n[k];
	double lim = (ni < __builtin_inf()) ? ni : __builtin_inf();
	double maxm = C.mu[k] * lim;
	double cm = sqrt(n0 * n0 + n1 * n1);
	double p0, p1;
	if (cm > maxm) {
		C.sready[k] = 0;
		double q = maxm / cm;
		p0 = n0 * q;
		p1 = n1 * q;
	} else {
		C.sready[k] = 1;
		p0 = n0;
		p1 = n1;
	}
	double e0 = p0 - i0, e1 = p1 - i1;
	if (!is_zero2(e0, e1)) {
		for (int r = 0; r < 6; r++) {
			double da = 0, db = 0;
			da += e0 * C.fita[(r * 2) * K + k];
			da += e1 * C.fita[(r * 2 + 1) * K + k];
			db += e0 * C.fitb[(r * 2) * K + k];
			db += e1 * C.fitb[(r * 2 + 1) * K + k];
			A.add(r, da);
			B.add(r, db);
		}
		C.imp_t[k] = i0 + e0;
		C.imp_t[K + k] = i1 + e1;
	}
}

__device__ __forceinline__ void apply_dv(Bodies& b, int i) {  // PhysicsEngine.cpp:1113-1136
	const int N = b.n;
	for (int c = 0; c < 3; c++) {
		b.vel[c * N + i] += b.dv[c * N + i];
		b.angv[c * N + i] += b.dv[(3 + c) * N + i];
		b.pvel[c * N + i] += b.dv[(6 + c) * N + i];
		b.pang[c * N + i] += b.dv[(9 + c) * N + i];
	}
}

// Sequential island solve: rows[2*r] = contact index, rows[2*r+1] = 0 normal / 1 friction.
__global__ void k_pgs_ordered(Bodies b, Contacts C, const int* rows, const int* isl_off, int nisl, Params P, int* applied) {
	int t = blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= nisl) return;
	int r0 = isl_off[t], r1 = isl_off[t + 1];
	auto refs = [&](int k, DvRef& A, DvRef& B) {
		int i1 = (int)(C.key[k] >> 32), i2 = (int)(C.key[k] & 0xFFFFFFFFu);
		A = DvRef{b.dv, i1, b.n, b.movement[i1] == 0};
		B = DvRef{b.dv, i2, b.n, b.movement[i2] == 0};
	};
	for (int r = r0; r < r1; r++) {
		int k = rows[2 * r];
		DvRef A, B;
		refs(k, A, B);
		const int K = C.cap;
		if (rows[2 * r + 1] == 0) {
			double imp = C.imp_n[k];
			if (!is_zero1(imp))
				for (int q = 0; q < 6; q++) {
					double da = 0, db = 0;
					da += imp * C.ita[q * K + k];
					db += imp * C.itb[q * K + k];
					A.add(q, da);
					B.add(q, db);
				}
		} else {
			double t0 = C.imp_t[k], t1 = C.imp_t[K + k];
			if (!is_zero2(t0, t1))
				for (int q = 0; q < 6; q++) {
					double da = 0, db = 0;
					da += t0 * C.fita[(q * 2) * K + k];
					da += t1 * C.fita[(q * 2 + 1) * K + k];
					db += t0 * C.fitb[(q * 2) * K + k];
					db += t1 * C.fitb[(q * 2 + 1) * K + k];
					A.add(q, da);
					B.add(q, db);
				}
		}
	}
	for (int it = 0; it < P.vel_iters; it++)
		for (int r = r0; r < r1; r++) {
			int k = rows[2 * r];
			DvRef A, B;
			refs(k, A, B);
			if (rows[2 * r + 1] == 0) normal_step(C, k, A, B, false);
			else friction_step(C, k, A, B);
		}
	for (int it = 0; it < P.pos_iters; it++)
		for (int r = r0; r < r1; r++) {
			if (rows[2 * r + 1] != 0) continue;
			int k = rows[2 * r];
			DvRef A, B;
			refs(k, A, B);
			normal_step(C, k, A, B, true);
		}
	for (int r = r0; r < r1; r++) {
		int k = rows[2 * r];
		int ii[2] = {(int)(C.key[k] >> 32), (int)(C.key[k] & 0xFFFFFFFFu)};
		for (int q = 0; q < 2; q++) {
			int i = ii[q];
			if (b.movement[i] != 0 || applied[i]) continue;
			applied[i] = 1;
			apply_dv(b, i);
		}
	}
}

// One workgroup per scene; colour-ordered sweeps.
__global__ void k_pgs_colored(Bodies b, Contacts C, const int* scene_body_off, const int* edge_range, const int* order,
                              const int* color_off, const int* ncolors, const int* efirst, const int* ecnt, Params P) {
	const int s = blockIdx.x;
	const int b0 = scene_body_off[s], b1 = scene_body_off[s + 1];
	const int N = b.n;
	for (int i = b0 + threadIdx.x; i < b1; i += blockDim.x)
		for (int c = 0; c < 12; c++) b.dv[c * N + i] = 0;
	__syncthreads();
	const int e0 = edge_range[2 * s];
	const int* co = color_off + (size_t)s * (MAXC + 2);
	const int nc = ncolors[s];
	// phase 0: warm start, 1: velocity sweeps, 2: pseudo-velocity sweeps
	int total = 1 + P.vel_iters + P.pos_iters;
	for (int ph = 0; ph < total; ph++) {
		int kind = ph == 0 ? 0 : (ph <= P.vel_iters ? 1 : 2);
		for (int c = 0; c < nc; c++) {
			int beg = co[c], end = co[c + 1];
			bool serial = (c == MAXC);
			int stride = serial ? 1 : blockDim.x;
			int start = serial ? (threadIdx.x == 0 ? beg : end) : beg + threadIdx.x;
			for (int x = start; x < end; x += stride) {
				int e = order[e0 + x];
				int k0 = efirst[e], k1 = k0 + ecnt[e];
				int i1 = (int)(C.key[k0] >> 32), i2 = (int)(C.key[k0] & 0xFFFFFFFFu);
				DvRef A{b.dv, i1, N, b.movement[i1] == 0}, B{b.dv, i2, N, b.movement[i2] == 0};
				for (int k = k0; k < k1; k++) {
					if (!C.live[k]) continue;
					if (kind == 0) warm_start(C, k, A, B);
					else if (kind == 1) {
						normal_step(C, k, A, B, false);
						friction_step(C, k, A, B);
					} else normal_step(C, k, A, B, true);
				}
			}
			__syncthreads();
		}
	}
	for (int i = b0 + threadIdx.x; i < b1; i += blockDim.x)
		if (b.movement[i] == 0) apply_dv(b, i);
}

// FP64 natural-map residual of the velocity rows after the sweeps; acc[0] +=
// sum r^2, acc[1] += sum lambda^2, acc[2] += rows.  Reads the final velocity
// deltas (b.dv) of the solve, so it runs right after the PGS launch.
__global__ void k_pgs_residual(Bodies b, Contacts C, int nc, double* acc) {
	int k = blockIdx.x * blockDim.x + threadIdx.x;
	double r2 = 0, l2 = 0, rows = 0;
	if (k < nc && C.live[k]) {
		const int K = C.cap, N = b.n;
		int i1 = (int)(C.key[k] >> 32), i2 = (int)(C.key[k] & 0xFFFFFFFFu);
		double x = 0, y = 0;
		for (int q = 0; q < 6; q++) {
			x += b.dv[q * N + i1] * C.ja[q * K + k];
			y += b.dv[q * N + i2] * C.jb[q * K + k];
		}
		double ln = C.imp_n[k];
		double z = ln + C.inv_n[k] * (C.tgt_n[k] - (x + y));
		double rn = ln - (z < 0 ? 0 : z);
		double cur[2];
		for (int r = 0; r < 2; r++) {
			double a = 0, c = 0;
			for (int q = 0; q < 6; q++) {
				a += b.dv[q * N + i1] * C.fja[(r * 6 + q) * K + k];
				c += b.dv[q * N + i2] * C.fjb[(r * 6 + q) * K + k];
			}
			cur[r] = a + c;
		}
		double d0 = C.ftgt[k] - cur[0], d1 = C.ftgt[K + k] - cur[1];
		double t0 = C.imp_t[k], t1 = C.imp_t[K + k];
		double z0 = t0 + C.finv[k] * d0 + C.finv[K + k] * d1;
		double z1 = t1 + C.finv[2 * K + k] * d0 + C.finv[3 * K + k] * d1;
		double lim = C.mu[k] * (ln > 0 ? ln : 0);
		double m = sqrt(z0 * z0 + z1 * z1);
		if (m > lim && m > 0) { z0 *= lim / m; z1 *= lim / m; }
		double r0 = t0 - z0, r1 = t1 - z1;
		r2 = rn * rn + r0 * r0 + r1 * r1;
		l2 = ln * ln + t0 * t0 + t1 * t1;
		rows = 3;
	}
	// wave reduction, one atomic per wave
	for (int off = 32; off >= 1; off >>= 1) {
		r2 += __shfl_xor(r2, off);
		l2 += __shfl_xor(l2, off);
		rows += __shfl_xor(rows, off);
	}
	if ((threadIdx.x & 63) == 0 && rows > 0) {
		atomicAdd(&acc[0], r2);
		atomicAdd(&acc[1], l2);
		atomicAdd(&acc[2], rows);
	}
}

static inline int nblk(int n, int t) { return (n + t - 1) / t; }

void launch_pgs_residual(const Bodies& b, const Contacts& C, int nc, double* acc, hipStream_t s) {
	if (nc) hipLaunchKernelGGL(k_pgs_residual, dim3(nblk(nc, 256)), dim3(256), 0, s, b, C, nc, acc);
}

void launch_build_rows(const Bodies& b, const Contacts& C, int nc, const Params& P, hipStream_t s) {
	if (nc) hipLaunchKernelGGL(k_build_rows, dim3(nblk(nc, 64)), dim3(64), 0, s, b, C, nc, P);
}
void launch_pgs_ordered(const Bodies& b, const Contacts& C, const int* rows, const int* isl_off, int nisl, const Params& P, int* applied,
                        hipStream_t s) {
	if (nisl) hipLaunchKernelGGL(k_pgs_ordered, dim3(nblk(nisl, 64)), dim3(64), 0, s, b, C, rows, isl_off, nisl, P, applied);
}
void launch_pgs_colored(const Bodies& b, const Contacts& C, const int* scene_body_off, int nscenes, const int* edge_range, const int* order,
                        const int* color_off, const int* ncolors, const int* efirst, const int* ecnt, const Params& P, int threads,
                        hipStream_t s) {
	if (nscenes)
		hipLaunchKernelGGL(k_pgs_colored, dim3(nscenes), dim3(threads), 0, s, b, C, scene_body_off, edge_range, order, color_off, ncolors,
		                   efirst, ecnt, P);
}

}  // namespace cpx
