// Constraint rows and projected Gauss-Seidel.
//
// k_build_rows restates Contact::updateSolverConstraints
// (PersistentConstraint.cpp:201-215) -> ContactConstraint (ConstraintSolver.cpp:
// 240-272) and FrictionConstraint (:277-343), plus the per-solve target values
// (:265-268 and ConstraintSolver.h:78).  Matrix products keep NMat's
// accumulation order (NMat.h:50-60, :91-103) so FP64 rows match bit for bit.
// Each contact's rows are one contiguous record of CP_RS values, so a row
// update reads a few whole cache lines instead of one line per component.
// Precision: FP64 records and deltas (parity), or the split path with FP32
// records, impulses and velocity deltas over the FP64 body state (rows are
// derived in FP64, deltas are added to FP64 velocities, integration is FP64).
//
// Sweep schedules over the same rows (ConstraintSolver.cpp:80-213):
//  * k_pgs_ordered: the reference's sequential order (one thread per island,
//    rows in the island-search order) — the parity mode.
//  * k_pgs_islands: the same islands with their edges graph-coloured, one wave
//    per island (joints and contacts; FP64).
//  * k_pgs_colored_lds / k_pgs_colored: one workgroup per scene; edges (body
//    pairs) are coloured so that no two edges of a colour share a dynamic
//    body, each thread runs one edge's contacts in cache order, and colours
//    are separated by a workgroup barrier.  Gauss-Seidel per body, in a
//    graph-coloured order.  The LDS variant keeps the scene's velocity deltas
//    in LDS; the global one serves scenes too large for LDS.
//
// Roofline: per contact per launch the algorithmic traffic is the row record
// (warm start, then per velocity sweep the normal and friction rows with
// their impulses, per position sweep the normal row), plus two bodies'
// velocity deltas when those are not in LDS.
#include "k_hsys.h"
#include "k_integrate.h"

namespace cpx {

#define MAXC CP_MAXC

__device__ __forceinline__ V3 ld3(const double* a, int i, int n) { return V3{a[i], a[n + i], a[2 * n + i]}; }

// Row record of contact k: CP_RS FP64 values, or the compact CP_RSF FP32 values
// of the split solve.
template <class T>
__host__ __device__ constexpr int rec_stride() { return sizeof(T) == 4 ? CP_RSF : CP_RS; }
template <class T>
__device__ __forceinline__ T* rec(const Contacts& C, int k) {
	return static_cast<T*>(__builtin_assume_aligned(reinterpret_cast<T*>(C.row) + (size_t)k * rec_stride<T>(), 128));
}

// The FP32 copy of body i's world inverse inertia, upper triangle (the split
// solve's rows multiply by it; 0 unless the body is dynamic, as getInvTensor
// returns for the others, RigidBody.cpp:169-171).
__device__ __forceinline__ void iinv32_body(const Bodies& b, int i) {
	const int N = b.n;
	const bool d = b.movement[i] == 0;
	auto e = [&](int r, int c) { return d ? (float)b.Iinv[(3 * r + c) * N + i] : 0.0f; };
	float4* o = reinterpret_cast<float4*>(b.iinv32 + 8 * (size_t)i);
	o[0] = float4{e(0, 0), e(0, 1), e(0, 2), e(1, 1)};
	o[1] = float4{e(1, 2), e(2, 2), 0.0f, 0.0f};
}
__global__ __launch_bounds__(256) void k_iinv32(Bodies b) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < b.nl) iinv32_body(b, i);
}
__device__ __forceinline__ void ld_iinv32(const float* base, int i, float (&m)[6]) {
	const float4* p = reinterpret_cast<const float4*>(base + 8 * (size_t)i);
	float4 x = p[0], y = p[1];
	m[0] = x.x; m[1] = x.y; m[2] = x.z; m[3] = x.w; m[4] = y.x; m[5] = y.y;
}

template <class T>
__device__ __forceinline__ T tsqrt(T x) {
	if constexpr (sizeof(T) == 4) return sqrtf(x);
	else return sqrt(x);
}

// Rows are derived in FP64 from the FP64 state and stored in the solve
// precision T (FP64, or FP32 for the split path).
// Contacts [0, n): n = *ncp when the count lives on the device (a captured
// graph replays with each step's count), else nc; grid-stride.
template <class T>
__device__ void build_row(const Bodies& b, const Contacts& C, int k, const Params& P);

// Each logical block (xcd_block) takes one contiguous chunk of contacts, so
// the contacts of a scene, which gather the same bodies, run on one XCD.
template <class T>
__global__ __launch_bounds__(64) void k_build_rows(Bodies b, Contacts C, int nc, const int* ncp, Params P) {
	const int n = ncp ? *ncp : nc;
	const int chunk = (((n + gridDim.x - 1) / gridDim.x) + 63) & ~63;
	const int k0 = xcd_block() * chunk, k1 = min(n, k0 + chunk);
	for (int k = k0 + threadIdx.x; k < k1; k += 64)
		if (C.live[k]) build_row<T>(b, C, k, P);
}

// The compact FP32 record of the split solve, derived in FP64 from the same
// state: the values the FP64 record's products give, evaluated directly
// (J M^-1 J^T as ima n.n + j_ang . I^-1 j_ang and its friction 2x2
// counterpart) instead of through 6x6 products, then rounded to FP32.  The
// inverse inertia is the bodies' FP32 copy (Bodies::iinv32, current when rows
// are built), the matrix the row steps multiply by.
__device__ void build_row_split(const Bodies& b, const Contacts& C, int k, const Params& P) {
	const int N = b.n, K = C.cap;
	float* R = rec<float>(C, k);
	const int i1 = (int)(C.key[k] >> 32), i2 = (int)(C.key[k] & 0xFFFFFFFFu);
	const Q q1{b.q[i1], b.q[N + i1], b.q[2 * N + i1], b.q[3 * N + i1]};
	const Q q2{b.q[i2], b.q[N + i2], b.q[2 * N + i2], b.q[3 * N + i2]};
	const V3 c1 = ld3(b.com, i1, N), c2 = ld3(b.com, i2, N);
	const V3 p1 = q_rotate(q1, ld3(C.lp1, k, K)) + c1;
	const V3 p2 = q_rotate(q2, ld3(C.lp2, k, K)) + c2;
	const V3 n = q_rotate(q1, ld3(C.ln1, k, K));
	const double adj = dot(p2 - p1, n) + C.pen[k];
	const double a1 = P.hardness * P.sub_dt, a2 = 1.0 / P.sub_dt;
	const double pcc = (a2 < a1) ? a2 : a1;  // posCorrectCoeff, PersistentConstraint.cpp:199
	const bool d1 = b.movement[i1] == 0, d2 = b.movement[i2] == 0;
	const double ima = d1 ? 1.0 / b.mass[i1] : 0, imb = d2 ? 1.0 / b.mass[i2] : 0;
	// the bodies' FP32 inverse inertia (upper triangles), the one the row steps multiply by
	float fa[6], fb[6];
	ld_iinv32(b.iinv32, i1, fa);
	ld_iinv32(b.iinv32, i2, fb);
	auto sym = [](const float (&m)[6], V3 v) {
		return V3{m[0] * v.x + m[1] * v.y + m[2] * v.z, m[1] * v.x + m[3] * v.y + m[4] * v.z, m[2] * v.x + m[4] * v.y + m[5] * v.z};
	};
	const V3 rA = p1 - c1, rB = p1 - c2;
	const bool x1 = b.movement[i1] == 1, x2 = b.movement[i2] == 1;  // getVel()/getAngVel() of FIXED bodies are zero
	const V3 z{0, 0, 0};
	const V3 va = x1 ? z : ld3(b.vel, i1, N), wa = x1 ? z : ld3(b.angv, i1, N);
	const V3 vb = x2 ? z : ld3(b.vel, i2, N), wb = x2 ? z : ld3(b.angv, i2, N);
	const double nn = dot(n, n);
	// normal row: J_a = (-n | n x rA), J_b = (n | rB x n)
	{
		const V3 ja = cross(n, rA), jb = cross(rB, n);
		double itv = ima * nn + dot(ja, sym(fa, ja)) + imb * nn + dot(jb, sym(fb, jb));
		itv *= (1 + P.cfm);
		R[CP_R_INV] = (float)(1.0 / itv);
		const double cur = (dot(va, -n) + dot(wa, ja)) + (dot(vb, n) + dot(wb, jb));
		R[CP_R_TGT] = (float)((cur < -P.cutoff_vel) ? -(1 + C.bounce[k]) * cur : -cur);
		double sp = adj - 0.05;
		sp = (sp < 0.0) ? 0.0 : sp;
		R[CP_R_PTGT] = (float)(sp * pcc);
		R[CP_R_PIMP] = 0;
		R[CP_R_IMP] = (float)C.imp_n[k];
		R[CP_R_MU] = (float)C.mu[k];
	}
	// friction rows: J = (u | rA x u), (-u | u x rB) and the same for w
	{
		const V3 su = ld3(C.fu, k, K), sw = ld3(C.fw, k, K);
		const double f0 = C.imp_t[k], f1 = C.imp_t[K + k];
		V3 u, w;
		perp_basis(n, &u, &w);
		const V3 au = cross(rA, u), aw = cross(rA, w), bu = cross(u, rB), bw = cross(w, rB);
		const V3 tau = sym(fa, au), taw = sym(fa, aw), tbu = sym(fb, bu), tbw = sym(fb, bw);
		const double m = ima + imb;
		double k00 = m * dot(u, u) + dot(au, tau) + dot(bu, tbu);
		const double k01 = m * dot(u, w) + dot(au, taw) + dot(bu, tbw);
		const double k10 = m * dot(w, u) + dot(aw, tau) + dot(bw, tbu);
		double k11 = m * dot(w, w) + dot(aw, taw) + dot(bw, tbw);
		k00 *= (1 + P.cfm);
		k11 *= (1 + P.cfm);
		const double di = 1.0 / (k00 * k11 - k01 * k10);
		R[CP_R_FINV] = (float)(k11 * di);
		R[CP_R_FINV + 1] = (float)(-k01 * di);
		R[CP_R_FINV + 2] = (float)(-k10 * di);
		R[CP_R_FINV + 3] = (float)(k00 * di);
		R[CP_R_FIMP] = (float)(dot(u, su) * f0 + dot(u, sw) * f1);
		R[CP_R_FIMP + 1] = (float)(dot(w, su) * f0 + dot(w, sw) * f1);
		R[CP_R_FTGT] = (float)(-((dot(va, u) + dot(wa, au)) + (dot(vb, -u) + dot(wb, bu))));
		R[CP_R_FTGT + 1] = (float)(-((dot(va, w) + dot(wa, aw)) + (dot(vb, -w) + dot(wb, bw))));
		R[CP_R_SREADY] = 0;
		C.fu[k] = u.x; C.fu[K + k] = u.y; C.fu[2 * K + k] = u.z;
		C.fw[k] = w.x; C.fw[K + k] = w.y; C.fw[2 * K + k] = w.z;
		C.sready[k] = 0;
		R[CP_F_U] = (float)u.x; R[CP_F_U + 1] = (float)u.y; R[CP_F_U + 2] = (float)u.z;
		R[CP_F_W] = (float)w.x; R[CP_F_W + 1] = (float)w.y; R[CP_F_W + 2] = (float)w.z;
	}
	R[CP_F_N] = (float)n.x; R[CP_F_N + 1] = (float)n.y; R[CP_F_N + 2] = (float)n.z;
	R[CP_F_IMA] = (float)ima;
	R[CP_F_IMB] = (float)imb;
	R[CP_F_RA] = (float)rA.x; R[CP_F_RA + 1] = (float)rA.y; R[CP_F_RA + 2] = (float)rA.z;
	R[CP_F_RB] = (float)rB.x; R[CP_F_RB + 1] = (float)rB.y; R[CP_F_RB + 2] = (float)rB.z;
}

template <class T>
__device__ void build_row(const Bodies& b, const Contacts& C, int k, const Params& P) {
	if constexpr (sizeof(T) == 4) {
		build_row_split(b, C, k, P);
		return;
	}
	const int N = b.n, K = C.cap;
	T* R = rec<T>(C, k);
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
	const double ima = d1 ? 1.0 / b.mass[i1] : 0, imb = d2 ? 1.0 / b.mass[i2] : 0;
	{
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
	// getVel()/getAngVel() are zero only for FIXED bodies (RigidBody.cpp:174-181)
	double va[6], vb[6];
	{
		bool x1 = b.movement[i1] == 1, x2 = b.movement[i2] == 1;
		V3 v = x1 ? V3{0, 0, 0} : ld3(b.vel, i1, N), w = x1 ? V3{0, 0, 0} : ld3(b.angv, i1, N);
		va[0] = v.x; va[1] = v.y; va[2] = v.z; va[3] = w.x; va[4] = w.y; va[5] = w.z;
		v = x2 ? V3{0, 0, 0} : ld3(b.vel, i2, N);
		w = x2 ? V3{0, 0, 0} : ld3(b.angv, i2, N);
		vb[0] = v.x; vb[1] = v.y; vb[2] = v.z; vb[3] = w.x; vb[4] = w.y; vb[5] = w.z;
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
		// stored: the basis and the angular parts (linear parts are -n | n and
		// ima, imb times those)
		R[CP_R_BASIS] = n.x; R[CP_R_BASIS + 1] = n.y; R[CP_R_BASIS + 2] = n.z; R[CP_R_BASIS + 3] = ima;
		R[CP_R_BASIS + 7] = imb;
		R[CP_R_BASIS + 11] = 0;
		for (int r = 0; r < 3; r++) {
			R[CP_R_ANG + r] = ja[3 + r]; R[CP_R_ANG + 3 + r] = jb[3 + r];
			R[CP_R_ANG + 6 + r] = ita[3 + r]; R[CP_R_ANG + 9 + r] = itb[3 + r];
		}
		R[CP_R_INV] = 1.0 / itv;
		double sp = adj - 0.05;
		sp = (sp < 0.0) ? 0.0 : sp;  // std::max<double>(0.0, x)
		R[CP_R_PTGT] = sp * pcc;
		R[CP_R_PIMP] = 0;
		R[CP_R_IMP] = cimp;
		R[CP_R_MU] = C.mu[k];
		double x = 0, y = 0;
		for (int q = 0; q < 6; q++) { x += va[q] * ja[q]; y += vb[q] * jb[q]; }
		double cur = x + y;
		R[CP_R_TGT] = (cur < -P.cutoff_vel) ? -(1 + C.bounce[k]) * cur : -cur;
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
		// rowMajorOrderInverse<2> (NMat.cpp:11-21)
		double a = itv[0][0], bq = itv[0][1], cq = itv[1][0], d = itv[1][1];
		double dt = a * d - bq * cq;
		double di = 1.0 / dt;
		R[CP_R_FINV] = d * di; R[CP_R_FINV + 1] = -bq * di; R[CP_R_FINV + 2] = -cq * di; R[CP_R_FINV + 3] = a * di;
		R[CP_R_BASIS + 4] = u.x; R[CP_R_BASIS + 5] = u.y; R[CP_R_BASIS + 6] = u.z;
		R[CP_R_BASIS + 8] = w.x; R[CP_R_BASIS + 9] = w.y; R[CP_R_BASIS + 10] = w.z;
		for (int r = 0; r < 2; r++)
			for (int c = 0; c < 3; c++) { R[CP_R_FANG + r * 3 + c] = ja[r][3 + c]; R[CP_R_FANG + 6 + r * 3 + c] = jb[r][3 + c]; }
		for (int r = 0; r < 3; r++)
			for (int c = 0; c < 2; c++) { R[CP_R_FANG + 12 + r * 2 + c] = ita[3 + r][c]; R[CP_R_FANG + 18 + r * 2 + c] = itb[3 + r][c]; }
		R[CP_R_FIMP] = i0;
		R[CP_R_FIMP + 1] = i1v;
		C.fu[k] = u.x; C.fu[K + k] = u.y; C.fu[2 * K + k] = u.z;
		C.fw[k] = w.x; C.fw[K + k] = w.y; C.fw[2 * K + k] = w.z;
		C.sready[k] = 0;
		R[CP_R_SREADY] = 0;
		for (int r = 0; r < 2; r++) {
			double x = 0, y = 0;
			for (int q = 0; q < 6; q++) { x += va[q] * ja[r][q]; y += vb[q] * jb[r][q]; }
			R[CP_R_FTGT + r] = -(x + y);
		}
	}
}

// ------------------------------------------------------------- row updates
// Velocity-delta accessors in the solve precision T: global memory [12][n]
// (components 0-5 velocity, 6-11 pseudo-velocity) or LDS [6][nbs] holding one
// half at a time.
template <class T>
struct DvGlobal {
	T* base;
	int i, n;
	bool dyn;
	__device__ T get(int c) const { return base[c * n + i]; }
	__device__ void add(int c, T v) const {
		if (dyn) base[c * n + i] += v;
	}
	__device__ void set(int c, T v) const { base[c * n + i] = v; }
};
template <class T>
struct DvLds {
	T* s;
	int i, nbs;  // scene-local body index
	bool dyn;
	__device__ T get(int c) const { return s[(c % 6) * nbs + i]; }
	__device__ void add(int c, T v) const {
		if (dyn) s[(c % 6) * nbs + i] += v;
	}
	__device__ void set(int c, T v) const { s[(c % 6) * nbs + i] = v; }
};
// One body's six deltas of the phase's half (velocity, or pseudo-velocity
// for position sweeps) held in registers while a thread runs the contacts of
// one coloured edge: no other thread touches a dynamic body of the edge during
// its colour, so the edge's contacts update the registers in the same order
// and rounding as they would the memory copy, which is written back once.
template <class T>
struct DvReg {
	mutable T v[6];
	bool dyn;
	__device__ T get(int c) const { return v[c % 6]; }
	__device__ void add(int c, T x) const {
		if (dyn) v[c % 6] += x;
	}
};

// Record blocks are loaded into registers as 16-byte vectors before any
// update is written, so a row step waits for memory once instead of once per
// coefficient (the deltas it writes are LDS or other bodies' memory, which the
// compiler cannot always prove disjoint from the record).
template <class T, int N>
__device__ __forceinline__ void ldv(const T* p, T (&o)[N]) {
	constexpr int W = 16 / sizeof(T);
	static_assert(N % W == 0, "whole 16-byte vectors");
	if constexpr (sizeof(T) == 4) {
		const float4* q = reinterpret_cast<const float4*>(p);
#pragma unroll
		for (int i = 0; i < N / 4; i++) {
			float4 x = q[i];
			o[4 * i] = x.x; o[4 * i + 1] = x.y; o[4 * i + 2] = x.z; o[4 * i + 3] = x.w;
		}
	} else {
		const double2* q = reinterpret_cast<const double2*>(p);
#pragma unroll
		for (int i = 0; i < N / 2; i++) {
			double2 x = q[i];
			o[2 * i] = x.x; o[2 * i + 1] = x.y;
		}
	}
}

// Register copies of a record: h = fields 0..15, j = J_a|J_b, it = ITA|ITB,
// fj = friction J_a|J_b, fit = friction ITA|ITB (blocks are adjacent).
template <class T>
struct RowRegs {
	T h[16], j[12], it[12], fj[24], fit[24];
};

// The normal row's J_a|J_b (g.j) and M^-1 J^T (g.it) from the basis block bs
// (n, 1/m_a, u, 1/m_b, w) and the angular block a.  The linear parts of M^-1
// J^T are 0 + (1/m) j: the full product's sum starts at +0 and adds five
// signed zeros, which leaves exactly this value.
template <class T>
__device__ __forceinline__ void expand_normal(const T (&bs)[12], const T (&a)[12], RowRegs<T>& g) {
#pragma unroll
	for (int q = 0; q < 3; q++) {
		g.j[q] = -bs[q];
		g.j[3 + q] = a[q];
		g.j[6 + q] = bs[q];
		g.j[9 + q] = a[3 + q];
		g.it[q] = T(0) + bs[3] * g.j[q];
		g.it[3 + q] = a[6 + q];
		g.it[6 + q] = T(0) + bs[7] * g.j[6 + q];
		g.it[9 + q] = a[9 + q];
	}
}
// The friction rows' J_a|J_b (g.fj: [dir][6] for a, then b) and M^-1 J^T
// (g.fit: [6][dir] for a, then b): J_a = (u | w), J_b = (-u | -w) in the
// linear part, angular parts from block f.
template <class T>
__device__ __forceinline__ void expand_friction(const T (&bs)[12], const T (&f)[24], RowRegs<T>& g) {
#pragma unroll
	for (int q = 0; q < 3; q++) {
		g.fj[q] = bs[4 + q];
		g.fj[6 + q] = bs[8 + q];
		g.fj[12 + q] = -bs[4 + q];
		g.fj[18 + q] = -bs[8 + q];
		g.fj[3 + q] = f[q];
		g.fj[9 + q] = f[3 + q];
		g.fj[15 + q] = f[6 + q];
		g.fj[21 + q] = f[9 + q];
#pragma unroll
		for (int c = 0; c < 2; c++) {
			g.fit[q * 2 + c] = T(0) + bs[3] * g.fj[c * 6 + q];
			g.fit[12 + q * 2 + c] = T(0) + bs[7] * g.fj[12 + c * 6 + q];
			g.fit[(3 + q) * 2 + c] = f[12 + q * 2 + c];
			g.fit[12 + (3 + q) * 2 + c] = f[18 + q * 2 + c];
		}
	}
}
template <class T>
__device__ __forceinline__ void load_normal(const T* R, const T (&bs)[12], RowRegs<T>& g) {
	T a[12];
	ldv(R + CP_R_ANG, a);
	expand_normal(bs, a, g);
}
template <class T>
__device__ __forceinline__ void load_friction(const T* R, const T (&bs)[12], RowRegs<T>& g) {
	T f[24];
	ldv(R + CP_R_FANG, f);
	expand_friction(bs, f, g);
}

// The same rows from the compact FP32 record r and the bodies' FP32 inverse
// inertia ia, ib (upper triangles): J_a = (-n | n x r_a), J_b = (n | r_b x n)
// for the normal row, J_a = (u | r_a x u), J_b = (-u | u x r_b) (and w) for
// the friction rows, the angular parts of M^-1 J^T the inverse inertia times
// those: the products the FP64 record stores, evaluated in FP32.
__device__ __forceinline__ void cross3(const float* a, const float* b, float* o) {
	o[0] = a[1] * b[2] - a[2] * b[1];
	o[1] = a[2] * b[0] - a[0] * b[2];
	o[2] = a[0] * b[1] - a[1] * b[0];
}
__device__ __forceinline__ void sym3(const float (&m)[6], const float* v, float* o) {
	o[0] = m[0] * v[0] + m[1] * v[1] + m[2] * v[2];
	o[1] = m[1] * v[0] + m[3] * v[1] + m[4] * v[2];
	o[2] = m[2] * v[0] + m[4] * v[1] + m[5] * v[2];
}
__device__ __forceinline__ void compact_normal(const float (&r)[CP_RSF], const float (&ia)[6], const float (&ib)[6], RowRegs<float>& g) {
	const float* n = r + CP_F_N;
	float ja[3], jb[3], ta[3], tb[3];
	cross3(n, r + CP_F_RA, ja);
	cross3(r + CP_F_RB, n, jb);
	sym3(ia, ja, ta);
	sym3(ib, jb, tb);
#pragma unroll
	for (int q = 0; q < 3; q++) {
		g.j[q] = -n[q];
		g.j[3 + q] = ja[q];
		g.j[6 + q] = n[q];
		g.j[9 + q] = jb[q];
		g.it[q] = 0.0f + r[CP_F_IMA] * g.j[q];
		g.it[3 + q] = ta[q];
		g.it[6 + q] = 0.0f + r[CP_F_IMB] * g.j[6 + q];
		g.it[9 + q] = tb[q];
	}
}
__device__ __forceinline__ void compact_friction(const float (&r)[CP_RSF], const float (&ia)[6], const float (&ib)[6], RowRegs<float>& g) {
	const float *u = r + CP_F_U, *w = r + CP_F_W, *ra = r + CP_F_RA, *rb = r + CP_F_RB;
	float au[3], aw[3], bu[3], bw[3], tau[3], taw[3], tbu[3], tbw[3];
	cross3(ra, u, au);
	cross3(ra, w, aw);
	cross3(u, rb, bu);
	cross3(w, rb, bw);
	sym3(ia, au, tau);
	sym3(ia, aw, taw);
	sym3(ib, bu, tbu);
	sym3(ib, bw, tbw);
#pragma unroll
	for (int q = 0; q < 3; q++) {
		g.fj[q] = u[q];
		g.fj[3 + q] = au[q];
		g.fj[6 + q] = w[q];
		g.fj[9 + q] = aw[q];
		g.fj[12 + q] = -u[q];
		g.fj[15 + q] = bu[q];
		g.fj[18 + q] = -w[q];
		g.fj[21 + q] = bw[q];
	}
#pragma unroll
	for (int q = 0; q < 3; q++) {
#pragma unroll
		for (int c = 0; c < 2; c++) {
			g.fit[q * 2 + c] = 0.0f + r[CP_F_IMA] * g.fj[c * 6 + q];
			g.fit[12 + q * 2 + c] = 0.0f + r[CP_F_IMB] * g.fj[12 + c * 6 + q];
		}
		g.fit[(3 + q) * 2] = tau[q];
		g.fit[(3 + q) * 2 + 1] = taw[q];
		g.fit[12 + (3 + q) * 2] = tbu[q];
		g.fit[12 + (3 + q) * 2 + 1] = tbw[q];
	}
}

// applyWarmStartVelocityChange of the contact row (kind 0) or the friction
// pair (kind 1) (ConstraintSolver.cpp:121-126).
template <class T, class D>
__device__ __forceinline__ void warm_normal(const RowRegs<T>& g, const D& A, const D& B) {
	T imp = g.h[CP_R_IMP];
	if (is_zero1(imp)) return;
#pragma unroll
	for (int r = 0; r < 6; r++) {
		T da = 0, db = 0;
		da += imp * g.it[r];
		db += imp * g.it[6 + r];
		A.add(r, da);
		B.add(r, db);
	}
}
template <class T, class D>
__device__ __forceinline__ void warm_friction(const RowRegs<T>& g, const D& A, const D& B) {
	T t0 = g.h[CP_R_FIMP], t1 = g.h[CP_R_FIMP + 1];
	if (is_zero2(t0, t1)) return;
#pragma unroll
	for (int r = 0; r < 6; r++) {
		T da = 0, db = 0;
		da += t0 * g.fit[r * 2];
		da += t1 * g.fit[r * 2 + 1];
		db += t0 * g.fit[12 + r * 2];
		db += t1 * g.fit[12 + r * 2 + 1];
		A.add(r, da);
		B.add(r, db);
	}
}

// PGSConstraintStep<1> for the contact row (ConstraintSolver.cpp:12-40, :270-272).
template <class T, class D>
__device__ __forceinline__ void normal_core(T* R, RowRegs<T>& g, const D& A, const D& B, bool pseudo) {
	int off = pseudo ? 6 : 0;
	T x = 0, y = 0;
#pragma unroll
	for (int q = 0; q < 6; q++) {
		x += A.get(off + q) * g.j[q];
		y += B.get(off + q) * g.j[6 + q];
	}
	T cur = x + y;
	T tgt = pseudo ? g.h[CP_R_PTGT] : g.h[CP_R_TGT];
	T acc = pseudo ? g.h[CP_R_PIMP] : g.h[CP_R_IMP];
	T delta = tgt - cur;
	T m = 0;
	m += delta * g.h[CP_R_INV];
	T nw = acc + m;
	T pr = (nw < 0) ? T(0) : nw;  // std::max<double>(nw, 0)
	T diff = pr - acc;
	if (!is_zero1(diff)) {
#pragma unroll
		for (int r = 0; r < 6; r++) {
			T da = 0, db = 0;
			da += diff * g.it[r];
			db += diff * g.it[6 + r];
			A.add(off + r, da);
			B.add(off + r, db);
		}
		T v = acc + diff;
		if (pseudo) {
			g.h[CP_R_PIMP] = v;
			R[CP_R_PIMP] = v;
		} else {
			g.h[CP_R_IMP] = v;
			R[CP_R_IMP] = v;
		}
	}
}

// PGSConstraintStep<2> with FrictionConstraint::projectValidImpulse (ConstraintSolver.cpp:331-343);
// the normal impulse is the one the contact row holds now.
template <class T, class D>
__device__ __forceinline__ void friction_core(T* R, RowRegs<T>& g, const D& A, const D& B) {
	T cur[2];
#pragma unroll
	for (int r = 0; r < 2; r++) {
		T x = 0, y = 0;
#pragma unroll
		for (int q = 0; q < 6; q++) {
			x += A.get(q) * g.fj[r * 6 + q];
			y += B.get(q) * g.fj[12 + r * 6 + q];
		}
		cur[r] = x + y;
	}
	T d0 = g.h[CP_R_FTGT] - cur[0], d1 = g.h[CP_R_FTGT + 1] - cur[1];
	T m0 = 0, m1 = 0;
	m0 += d0 * g.h[CP_R_FINV];
	m0 += d1 * g.h[CP_R_FINV + 1];
	m1 += d0 * g.h[CP_R_FINV + 2];
	m1 += d1 * g.h[CP_R_FINV + 3];
	T i0 = g.h[CP_R_FIMP], i1 = g.h[CP_R_FIMP + 1];
	T n0 = i0 + m0, n1 = i1 + m1;
	T ni = g.h[CP_R_IMP];
	const T inf = (T)__builtin_inf();
	T lim = (ni < inf) ? ni : inf;
	T maxm = g.h[CP_R_MU] * lim;
	T cm = tsqrt(n0 * n0 + n1 * n1);
	T p0, p1;
	if (cm > maxm) {
		R[CP_R_SREADY] = 0;
		T q = maxm / cm;
		p0 = n0 * q;
		p1 = n1 * q;
	} else {
		R[CP_R_SREADY] = 1;
		p0 = n0;
		p1 = n1;
	}
	T e0 = p0 - i0, e1 = p1 - i1;
	if (!is_zero2(e0, e1)) {
#pragma unroll
		for (int r = 0; r < 6; r++) {
			T da = 0, db = 0;
			da += e0 * g.fit[r * 2];
			da += e1 * g.fit[r * 2 + 1];
			db += e0 * g.fit[12 + r * 2];
			db += e1 * g.fit[12 + r * 2 + 1];
			A.add(r, da);
			B.add(r, db);
		}
		g.h[CP_R_FIMP] = i0 + e0;
		g.h[CP_R_FIMP + 1] = i1 + e1;
		R[CP_R_FIMP] = i0 + e0;
		R[CP_R_FIMP + 1] = i1 + e1;
	}
}

// One contact in one phase: 0 warm start (both rows), 1 velocity sweep
// (normal row, then friction rows), 2 position sweep (normal row).  ia, ib:
// the bodies' FP32 inverse inertia (the compact FP32 record; unused in FP64).
// The compact FP32 record already in registers (r); R receives the updates.
// The row steps of warm_*, normal_core and friction_core with the rows
// formed as they are used (compact_normal / compact_friction), so that no
// expanded copy of the record stays live.
template <class D>
__device__ __forceinline__ void contact_phase_f32(float* R, const float (&r)[CP_RSF], const D& A, const D& B, int phase, const float (&ia)[6],
                                                  const float (&ib)[6]) {
	const float* n = r + CP_F_N;
	const float ima = r[CP_F_IMA], imb = r[CP_F_IMB];
	float ja[3], jb[3], ta[3], tb[3];
	cross3(n, r + CP_F_RA, ja);
	cross3(r + CP_F_RB, n, jb);
	sym3(ia, ja, ta);
	sym3(ib, jb, tb);
	// impulse x along the normal row into the deltas of half off
	auto push_normal = [&](float x, int off) {
#pragma unroll
		for (int q = 0; q < 3; q++) {
			float da = 0, db = 0;
			da += x * (0.0f + ima * -n[q]);
			db += x * (0.0f + imb * n[q]);
			A.add(off + q, da);
			B.add(off + q, db);
		}
#pragma unroll
		for (int q = 0; q < 3; q++) {
			float da = 0, db = 0;
			da += x * ta[q];
			db += x * tb[q];
			A.add(off + 3 + q, da);
			B.add(off + 3 + q, db);
		}
	};
	// returns the row's accumulated impulse after the step
	auto normal_step = [&](bool pseudo) -> float {
		const int off = pseudo ? 6 : 0;
		float x = 0, y = 0;
#pragma unroll
		for (int q = 0; q < 3; q++) {
			x += A.get(off + q) * -n[q];
			y += B.get(off + q) * n[q];
		}
#pragma unroll
		for (int q = 0; q < 3; q++) {
			x += A.get(off + 3 + q) * ja[q];
			y += B.get(off + 3 + q) * jb[q];
		}
		const float cur = x + y;
		const float tgt = pseudo ? r[CP_R_PTGT] : r[CP_R_TGT];
		const float acc = pseudo ? r[CP_R_PIMP] : r[CP_R_IMP];
		const float delta = tgt - cur;
		float m = 0;
		m += delta * r[CP_R_INV];
		const float nw = acc + m;
		const float pr = (nw < 0) ? 0.0f : nw;
		const float diff = pr - acc;
		if (is_zero1(diff)) return acc;
		push_normal(diff, off);
		const float v = acc + diff;
		R[pseudo ? CP_R_PIMP : CP_R_IMP] = v;
		return v;
	};
	if (phase == 2) {
		(void)normal_step(true);
		return;
	}
	const float *u = r + CP_F_U, *w = r + CP_F_W, *ra = r + CP_F_RA, *rb = r + CP_F_RB;
	float au[3], aw[3], bu[3], bw[3];
	cross3(ra, u, au);
	cross3(ra, w, aw);
	cross3(u, rb, bu);
	cross3(w, rb, bw);
	float tau[3], taw[3], tbu[3], tbw[3];
	sym3(ia, au, tau);
	sym3(ia, aw, taw);
	sym3(ib, bu, tbu);
	sym3(ib, bw, tbw);
	auto push_friction = [&](float t0, float t1) {
#pragma unroll
		for (int q = 0; q < 3; q++) {
			float da = 0, db = 0;
			da += t0 * (0.0f + ima * u[q]);
			da += t1 * (0.0f + ima * w[q]);
			db += t0 * (0.0f + imb * -u[q]);
			db += t1 * (0.0f + imb * -w[q]);
			A.add(q, da);
			B.add(q, db);
		}
#pragma unroll
		for (int q = 0; q < 3; q++) {
			float da = 0, db = 0;
			da += t0 * tau[q];
			da += t1 * taw[q];
			db += t0 * tbu[q];
			db += t1 * tbw[q];
			A.add(3 + q, da);
			B.add(3 + q, db);
		}
	};
	if (phase == 0) {  // warm start: the stored impulses (applyWarmStartVelocityChange)
		const float imp = r[CP_R_IMP];
		if (!is_zero1(imp)) push_normal(imp, 0);
		const float t0 = r[CP_R_FIMP], t1 = r[CP_R_FIMP + 1];
		if (!is_zero2(t0, t1)) push_friction(t0, t1);
		return;
	}
	const float ni = normal_step(false);
	// friction pair, projected into the cone of the normal impulse held now
	float c0x = 0, c0y = 0, c1x = 0, c1y = 0;
#pragma unroll
	for (int q = 0; q < 3; q++) {
		c0x += A.get(q) * u[q];
		c0y += B.get(q) * -u[q];
		c1x += A.get(q) * w[q];
		c1y += B.get(q) * -w[q];
	}
#pragma unroll
	for (int q = 0; q < 3; q++) {
		c0x += A.get(3 + q) * au[q];
		c0y += B.get(3 + q) * bu[q];
		c1x += A.get(3 + q) * aw[q];
		c1y += B.get(3 + q) * bw[q];
	}
	const float d0 = r[CP_R_FTGT] - (c0x + c0y), d1 = r[CP_R_FTGT + 1] - (c1x + c1y);
	float m0 = 0, m1 = 0;
	m0 += d0 * r[CP_R_FINV];
	m0 += d1 * r[CP_R_FINV + 1];
	m1 += d0 * r[CP_R_FINV + 2];
	m1 += d1 * r[CP_R_FINV + 3];
	const float i0 = r[CP_R_FIMP], i1 = r[CP_R_FIMP + 1];
	const float n0 = i0 + m0, n1 = i1 + m1;
	const float inf = __builtin_inff();
	const float lim = (ni < inf) ? ni : inf;
	const float maxm = r[CP_R_MU] * lim;
	const float cm = sqrtf(n0 * n0 + n1 * n1);
	float p0, p1;
	if (cm > maxm) {
		R[CP_R_SREADY] = 0;
		const float q = maxm / cm;
		p0 = n0 * q;
		p1 = n1 * q;
	} else {
		R[CP_R_SREADY] = 1;
		p0 = n0;
		p1 = n1;
	}
	const float e0 = p0 - i0, e1 = p1 - i1;
	if (!is_zero2(e0, e1)) {
		push_friction(e0, e1);
		R[CP_R_FIMP] = i0 + e0;
		R[CP_R_FIMP + 1] = i1 + e1;
	}
}
template <class T, class D>
__device__ __forceinline__ void contact_phase(T* R, const D& A, const D& B, int phase, const float (&ia)[6], const float (&ib)[6]) {
	RowRegs<T> g;
	if constexpr (sizeof(T) == 4) {
		float r[CP_RSF];
		ldv(R, r);
		contact_phase_f32(R, r, A, B, phase, ia, ib);
		return;
	} else {
		T bs[12];
		ldv(R, g.h);
		ldv(R + CP_R_BASIS, bs);
		load_normal(R, bs, g);
		if (phase != 2) load_friction(R, bs, g);
	}
	if (phase == 0) {
		warm_normal(g, A, B);
		warm_friction(g, A, B);
	} else if (phase == 1) {
		normal_core(R, g, A, B, false);
		friction_core(R, g, A, B);
	} else {
		normal_core(R, g, A, B, true);
	}
}

// One row (kind 0 contact normal, 1 friction) of an island's sequential list.
template <class T, class D>
__device__ __forceinline__ void contact_row(T* R, const D& A, const D& B, int kind, int phase) {
	RowRegs<T> g;
	T bs[12];
	ldv(R, g.h);
	ldv(R + CP_R_BASIS, bs);
	if (kind == 0) {
		load_normal(R, bs, g);
		if (phase == 0) warm_normal(g, A, B);
		else normal_core(R, g, A, B, phase == 2);
	} else {
		load_friction(R, bs, g);
		if (phase == 0) warm_friction(g, A, B);
		else friction_core(R, g, A, B);
	}
}

// applyVelocityChange (PhysicsEngine.cpp:1113-1136): velocity deltas in T
// added to the FP64 velocities.
template <class T>
__device__ __forceinline__ void apply_dv(Bodies& b, const T* dv, int i) {
	const int N = b.n;
	for (int c = 0; c < 3; c++) {
		b.vel[c * N + i] += (double)dv[c * N + i];
		b.angv[c * N + i] += (double)dv[(3 + c) * N + i];
		b.pvel[c * N + i] += (double)dv[(6 + c) * N + i];
		b.pang[c * N + i] += (double)dv[(9 + c) * N + i];
	}
}

// Island solves (FP64 only) over the island schedule: rows[2*r] =
// index, rows[2*r+1] = kind (0 contact normal, 1 friction, 2 joint).  At the
// first substep each island's holonomic systems are solved after the sweeps
// (ConstraintSolver.cpp:163-196).  The reference keeps one set of velocity
// changes per island solve (ConstraintSolver.cpp:83-114); dynamic bodies
// belong to one island, so their deltas live in b.dv, and the
// linearization-error vectors of non-dynamic joint bodies get one column per
// (island, body) (S.jlin).
using DG = DvGlobal<double>;

// Where an island solve keeps its velocity deltas: b.dv ([12][n], global body
// indices) or an island-local LDS array ([12][stride], local indices from the
// island search's cloc / jloc maps; non-dynamic bodies share the never-written zero
// column stride - 1).
struct DvMap {
	double* base;
	int stride;
	const int* cloc;  // contact k -> local (body 1, body 2); null: global indices
	const int* jloc;  // joint k -> local (a, b)
	int split;        // contact rows are the split solve's compact FP32 records
};

// An FP64 delta accessor seen in FP32 by the split contact rows: the row
// reads the shared deltas rounded to FP32 and adds its FP32 changes back.
template <class D>
struct DvAsF32 {
	const D& d;
	bool dyn;
	__device__ float get(int c) const { return (float)d.get(c); }
	__device__ void add(int c, float v) const { d.add(c, (double)v); }
};

__device__ __forceinline__ void contact_refs(const Bodies& b, const Contacts& C, const DvMap& m, int k, DG& A, DG& B) {
	int i1 = (int)(C.key[k] >> 32), i2 = (int)(C.key[k] & 0xFFFFFFFFu);
	int l1 = m.cloc ? m.cloc[2 * k] : i1, l2 = m.cloc ? m.cloc[2 * k + 1] : i2;
	A = DG{m.base, l1, m.stride, b.movement[i1] == 0};
	B = DG{m.base, l2, m.stride, b.movement[i2] == 0};
}
__device__ __forceinline__ int jl(const Joints& J, const DvMap& m, int k, int h) {
	return m.jloc ? m.jloc[2 * k + h] : (h ? J.b[k] : J.a[k]);
}

// Joint target at solve time (findTargetConstraintValue, ConstraintSolver.h:78)
// from getVel / getAngVel (RigidBody.cpp:174-181).
__device__ __forceinline__ void joint_target(const Bodies& b, const Joints& J, int k) {
	const int N = b.n;
	double* R = J.row + (size_t)k * CP_JRS;
	int ia = J.a[k], ib = J.b[k];
	double va[6], vb[6], val[CP_JMAX];
	bool fa = b.movement[ia] == 1, fb = b.movement[ib] == 1;
	for (int c = 0; c < 3; c++) {
		va[c] = fa ? 0.0 : b.vel[c * N + ia];
		va[3 + c] = fa ? 0.0 : b.angv[c * N + ia];
		vb[c] = fb ? 0.0 : b.vel[c * N + ib];
		vb[3 + c] = fb ? 0.0 : b.angv[c * N + ib];
	}
	joint_value(R, J.type[k], va, vb, 1, val);
	if (J.kind[k] == 10 || J.kind[k] == 11) R[CP_J_TGT] = R[CP_J_TV] - val[0];  // Motor / PistonConstraint::findTargetConstraintValue
	else
		for (int c = 0; c < J.type[k]; c++) R[CP_J_TGT + c] = -val[c];
}

// PGSConstraintStep<n> of a joint (ConstraintSolver.cpp:12-40); the limits
// project the accumulated impulse.
__device__ __forceinline__ void joint_step(const Bodies& b, const Joints& J, const DvMap& m, int k, bool pseudo) {
	const int K = J.n, S = m.stride;
	double* R = J.row + (size_t)k * CP_JRS;
	int n = J.type[k], ia = J.a[k], ib = J.b[k];
	double* va = m.base + (pseudo ? 6 * S : 0) + jl(J, m, k, 0);
	double* vb = m.base + (pseudo ? 6 * S : 0) + jl(J, m, k, 1);
	double* acc = pseudo ? J.pimp : J.imp;
	const double* tgt = R + (pseudo ? CP_J_PTGT : CP_J_TGT);
	double cur[CP_JMAX], delta[CP_JMAX], mm[CP_JMAX], diff[CP_JMAX], a[CP_JMAX];
	joint_value(R, n, va, vb, S, cur);
	for (int c = 0; c < n; c++) {
		delta[c] = tgt[c] - cur[c];
		a[c] = acc[c * K + k];
	}
	for (int i = 0; i < n; i++) {
		mm[i] = 0;
		for (int c = 0; c < n; c++) mm[i] += delta[c] * R[CP_J_INV + i * n + c];
	}
	for (int c = 0; c < n; c++) diff[c] = (a[c] + mm[c]) - a[c];
	const int pj = J.proj[k] & 3;
	if (pj) {  // projectValidImpulse of the inequality rows (ConstraintSolver.cpp:432-775), on the accumulation
		double nw = a[0] + mm[0], pr;
		// std::min / std::max<double>: signed zeros and NaN as they pass
		if (pj == CP_PROJ_CLAMP) {
			const double bd = R[CP_J_BOUND];
			pr = nw > 0 ? (bd < nw ? bd : nw) : (nw < -bd ? -bd : nw);
		} else {
			pr = pj == CP_PROJ_UPPER ? (0.0 < nw ? 0.0 : nw) : (nw < 0.0 ? 0.0 : nw);
		}
		diff[0] = pr - a[0];
	}
	if (!is_zero_n(diff, n)) {
		joint_apply(R, n, diff, va, vb, S, b.movement[ia] == 0, b.movement[ib] == 0);
		for (int c = 0; c < n; c++) acc[c * K + k] = a[c] + diff[c];
	}
}

// One row of a phase: 0 warm start, 1 velocity sweep, 2 position sweep.
__device__ __forceinline__ void island_row(const Bodies& b, const Contacts& C, const Joints& J, const DvMap& m, int k, int kind, int phase) {
	if (kind == 2) {
		if (phase == 0) {
			const int K = J.n;
			double imp[CP_JMAX];
			int n = J.type[k];
			for (int c = 0; c < n; c++) imp[c] = J.imp[c * K + k];
			if (!is_zero_n(imp, n))
				joint_apply(J.row + (size_t)k * CP_JRS, n, imp, m.base + jl(J, m, k, 0), m.base + jl(J, m, k, 1), m.stride,
				            b.movement[J.a[k]] == 0, b.movement[J.b[k]] == 0);
		} else if (phase == 1 || !(J.proj[k] & CP_PROJ_NOPOS)) {  // needsPosCorrect (ConstraintSolver.cpp:152)
			joint_step(b, J, m, k, phase == 2);
		}
		return;
	}
	if (phase == 2 && kind != 0) return;
	DG A, B;
	contact_refs(b, C, m, k, A, B);
	if (m.split) {
		// a contact's normal and friction rows are adjacent in every island row
		// list (Contact::addSolverConstraints): both run at the normal row
		if (kind != 0) return;
		float ia[6], ib[6];
		ld_iinv32(b.iinv32, (int)(C.key[k] >> 32), ia);
		ld_iinv32(b.iinv32, (int)(C.key[k] & 0xFFFFFFFFu), ib);
		const DvAsF32<DG> fa{A, A.dyn}, fb{B, B.dyn};
		float r[CP_RSF];
		ldv(rec<float>(C, k), r);
		contact_phase_f32(rec<float>(C, k), r, fa, fb, phase, ia, ib);
		return;
	}
	contact_row(rec<double>(C, k), A, B, kind, phase);
}

// computeAndApplyImpulses (HolonomicBlockSolver.cpp:298-355) for the island's
// systems in order, hol_iters times.
__device__ void island_holonomic(const Bodies& b, const Joints& J, const HSysDev& H, const IslandSched& S, int t, const Params& P,
                                 double* lin, const DvMap& m) {
	const int K = J.n, NL = S.nl, DS = m.stride;
	auto apply = [&](int s, bool pseudo, double thr) {
		const int nc = H.nc[s];
		const int* ord = H.order + H.ord_off[s];
		const int* vl = H.vloc + H.ord_off[s];
		int deg = H.degree[s];
		double* d = H.scratch + H.dbuf_off[s] / 2;  // 3 * degree doubles per system
		double* eff = d + deg;
		double* tmp = eff + deg;
		double* dv = m.base + (pseudo ? 6 * DS : 0);
		for (int i = 0; i < nc; i++) {
			int k = ord[i];
			const double* R = J.row + (size_t)k * CP_JRS;
			double cur[CP_JMAX];
			joint_value(R, J.type[k], dv + jl(J, m, k, 0), dv + jl(J, m, k, 1), DS, cur);
			const double* tg = R + (pseudo ? CP_J_PTGT : CP_J_TGT);
			for (int c = 0; c < J.type[k]; c++) d[vl[i] + c] = tg[c] - cur[c];
		}
		hs_solve(J, H, s, d, eff, tmp);
		if (thr > 0) {
			for (int i = 0; i < nc; i++) {
				int k = ord[i];
				if (is_zero_n(d + vl[i], J.type[k])) continue;
				joint_apply(J.row + (size_t)k * CP_JRS, J.type[k], d + vl[i], lin + S.jlin[2 * k], lin + S.jlin[2 * k + 1], NL, true, true);
			}
			for (int i = 0; i < nc; i++) {
				int k = ord[i];
				double r2 = 0;
				for (int c = 3; c < 6; c++) {
					double la = lin[c * NL + S.jlin[2 * k]], lb = lin[c * NL + S.jlin[2 * k + 1]];
					r2 += la * la + lb * lb;
				}
				if (r2 > thr * thr) return;  // discard the solution
			}
		}
		for (int i = 0; i < nc; i++) {
			int k = ord[i];
			int n = J.type[k];
			if (is_zero_n(d + vl[i], n)) continue;
			joint_apply(J.row + (size_t)k * CP_JRS, n, d + vl[i], dv + jl(J, m, k, 0), dv + jl(J, m, k, 1), DS, b.movement[J.a[k]] == 0,
			            b.movement[J.b[k]] == 0);
			double* acc = pseudo ? J.pimp : J.himp;
			for (int c = 0; c < n; c++) acc[c * K + k] += d[vl[i] + c];
		}
	};
	for (int it = 0; it < P.hol_iters; it++)
		for (int q = S.sys_off[t]; q < S.sys_off[t] + S.sys_cnt[t]; q++) {
			apply(S.sys[q], false, -1.0);
			apply(S.sys[q], true, (double)15.0f);  // excessive_linear_error_torque_threshold (PhysicsEngine.h:217)
		}
}

// addScaledHolonomicImpulseToNextWarmStart with factor 0 (PhysicsEngine.cpp:373).
__device__ __forceinline__ void holonomic_warm_scale(const Joints& J, int k) {
	if (J.sys[k] < 0) return;
	for (int c = 0; c < J.type[k]; c++) J.imp[c * J.n + k] += (double)0.0f * J.himp[c * J.n + k];
}

__device__ __forceinline__ void row_bodies(const Contacts& C, const Joints& J, int k, int kind, int* ii) {
	if (kind == 2) {
		ii[0] = J.a[k];
		ii[1] = J.b[k];
	} else {
		ii[0] = (int)(C.key[k] >> 32);
		ii[1] = (int)(C.key[k] & 0xFFFFFFFFu);
	}
}

// After an island's velocity changes (applyVelocityChange, PhysicsEngine.cpp:
// 1113-1136): an asleep body the solve moved faster than wake_vel wakes the
// whole island (wakeupIsland, :1285-1290; RigidBody::wake, RigidBody.cpp:
// 331-338); otherwise asleep bodies keep a zero velocity.  The reference
// visits the island's bodies in the order of an unordered_map keyed by
// address (ConstraintSolver.cpp:91, :199-212), so an asleep body it visits
// before a waking one is zeroed first; here every waking body comes first,
// one of the orders the reference can take.
__device__ __forceinline__ bool wakes(const Bodies& b, double wake_vel, int i) {
	const int N = b.n;
	V3 v{b.vel[i], b.vel[N + i], b.vel[2 * N + i]}, w{b.angv[i], b.angv[N + i], b.angv[2 * N + i]};
	return b.asleep[i] && (mag(v) > wake_vel || mag(w) > wake_vel);
}
__device__ __forceinline__ void wake_body(const Bodies& b, const SleepDev& sl, int i) {
	if (!b.asleep[i]) return;
	b.asleep[i] = 0;
	sl.counter[i] = 0;
	sl.ticks[i] = 0;
	sl.hlen[i] = 0;
	sl.hhead[i] = 0;
}
__device__ __forceinline__ void zero_asleep(const Bodies& b, int i) {
	if (!b.asleep[i]) return;
	const int N = b.n;
	for (int c = 0; c < 3; c++) {
		b.vel[c * N + i] = 0;
		b.angv[c * N + i] = 0;
	}
}

// Reference order: one thread per island walks its rows sequentially.
__global__ __launch_bounds__(64) void k_pgs_ordered(Bodies b, Contacts C, Joints J, HSysDev H, IslandSched S, Params P, int* applied, double* lin, int first) {
	int t = blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= S.n) return;
	const int* rows = S.rows;
	int r0 = S.off[t], r1 = S.off[t + 1];
	const DvMap m{b.dv, b.n, nullptr, nullptr};
	for (int r = r0; r < r1; r++)
		if (rows[2 * r + 1] == 2) joint_target(b, J, rows[2 * r]);
	for (int r = r0; r < r1; r++) island_row(b, C, J, m, rows[2 * r], rows[2 * r + 1], 0);
	for (int it = 0; it < P.vel_iters; it++)
		for (int r = r0; r < r1; r++) island_row(b, C, J, m, rows[2 * r], rows[2 * r + 1], 1);
	for (int it = 0; it < P.pos_iters; it++)
		for (int r = r0; r < r1; r++) island_row(b, C, J, m, rows[2 * r], rows[2 * r + 1], 2);
	if (first && P.hol_iters > 0 && S.sys_cnt[t] > 0) {
		island_holonomic(b, J, H, S, t, P, lin, m);
		for (int r = r0; r < r1; r++)
			if (rows[2 * r + 1] == 2) holonomic_warm_scale(J, rows[2 * r]);
	}
	for (int r = r0; r < r1; r++) {
		int ii[2];
		row_bodies(C, J, rows[2 * r], rows[2 * r + 1], ii);
		for (int q = 0; q < 2; q++) {
			int i = ii[q];
			if (b.movement[i] != 0 || applied[i]) continue;
			applied[i] = 1;
			apply_dv(b, b.dv, i);
		}
	}
	if (!b.asleep) return;
	bool wake = false;
	for (int r = r0; r < r1 && !wake; r++) {
		int ii[2];
		row_bodies(C, J, rows[2 * r], rows[2 * r + 1], ii);
		for (int q = 0; q < 2; q++)
			if (b.movement[ii[q]] == 0 && wakes(b, S.sl.wake_vel, ii[q])) wake = true;
	}
	if (wake) {
		for (int x = S.voff[t]; x < S.voff[t] + S.vcnt[t]; x++) wake_body(b, S.sl, S.vlist[x]);
		return;
	}
	for (int r = r0; r < r1; r++) {
		int ii[2];
		row_bodies(C, J, rows[2 * r], rows[2 * r + 1], ii);
		for (int q = 0; q < 2; q++)
			if (b.movement[ii[q]] == 0) zero_asleep(b, ii[q]);
	}
}

// Coloured islands: one wave per island.  The island's edges (body pairs with
// rows: joints, then live contacts) are coloured by the island search so that no two
// edges of a colour share a dynamic body; a lane runs one edge's rows in
// order and colours are separated by a workgroup barrier.  Per body the
// deltas accumulate in the order of the island's row list (colour-major), so
// the solve equals the sequential sweep over that list.  The island's last
// colour may be a serial bucket (edges that found no colour), run by lane 0.
__global__ __launch_bounds__(64) void k_pgs_islands(Bodies b, Contacts C, Joints J, HSysDev H, IslandSched S, Params P, int* applied,
                                                     double* lin, int first) {
	extern __shared__ double sdv[];
	const int t = blockIdx.x, lane = threadIdx.x;
	const int* rows = S.rows;
	const int r0 = S.off[t], r1 = S.off[t + 1];
	const int c0 = S.icoff[t], c1 = c0 + S.ncol[t];
	const bool ser = S.serial[t] != 0;
	// island-local velocity deltas in LDS when the island fits (its bodies
	// plus a zero column for the non-dynamic ones), else b.dv
	const int bo0 = S.boff[t], nbi = S.nbod[t];
	const bool local = S.cloc && nbi + 1 <= S.lds_bodies;
	const int split = P.precision == CP_PREC_SPLIT;
	const DvMap m = local ? DvMap{sdv, nbi + 1, S.cloc, S.jloc, split} : DvMap{b.dv, b.n, nullptr, nullptr, split};
	if (local)
		for (int x = lane; x < 12 * (nbi + 1); x += 64) sdv[x] = 0;
	for (int r = r0 + lane; r < r1; r += 64)
		if (rows[2 * r + 1] == 2) joint_target(b, J, rows[2 * r]);
	__syncthreads();
	auto sweep = [&](int phase) {
		for (int c = c0; c < c1; c++) {
			int e0 = S.coff[c], e1 = S.coff[c + 1];
			bool serial = ser && c == c1 - 1;
			int e = serial ? (lane == 0 ? e0 : e1) : e0 + lane;
			for (; e < e1; e += serial ? 1 : 64)
				for (int r = S.eoff[e]; r < S.eoff[e + 1]; r++) island_row(b, C, J, m, rows[2 * r], rows[2 * r + 1], phase);
			__syncthreads();
		}
	};
	sweep(0);
	for (int it = 0; it < P.vel_iters; it++) sweep(1);
	for (int it = 0; it < P.pos_iters; it++) sweep(2);
	if (first && P.hol_iters > 0 && S.sys_cnt[t] > 0) {
		if (lane == 0) island_holonomic(b, J, H, S, t, P, lin, m);
		__syncthreads();
		for (int r = r0 + lane; r < r1; r += 64)
			if (rows[2 * r + 1] == 2) holonomic_warm_scale(J, rows[2 * r]);
	}
	if (local) {
		// applyVelocityChange (PhysicsEngine.cpp:1113-1136) for the island's dynamic bodies
		const int N = b.n, DS = nbi + 1;
		for (int j = lane; j < nbi; j += 64) {
			int i = S.bod[bo0 + j];
			for (int c = 0; c < 3; c++) {
				b.vel[c * N + i] += sdv[c * DS + j];
				b.angv[c * N + i] += sdv[(3 + c) * DS + j];
				b.pvel[c * N + i] += sdv[(6 + c) * DS + j];
				b.pang[c * N + i] += sdv[(9 + c) * DS + j];
			}
		}
		if (!b.asleep) return;
		__syncthreads();
		bool w = false;
		for (int j = lane; j < nbi; j += 64) w = w || wakes(b, S.sl.wake_vel, S.bod[bo0 + j]);
		if (__syncthreads_or(w)) {
			for (int x = S.voff[t] + lane; x < S.voff[t] + S.vcnt[t]; x += 64) wake_body(b, S.sl, S.vlist[x]);
		} else {
			for (int j = lane; j < nbi; j += 64) zero_asleep(b, S.bod[bo0 + j]);
		}
		return;
	}
	for (int r = r0 + lane; r < r1; r += 64) {
		int ii[2];
		row_bodies(C, J, rows[2 * r], rows[2 * r + 1], ii);
		for (int q = 0; q < 2; q++) {
			int i = ii[q];
			if (b.movement[i] != 0 || atomicExch(applied + i, 1) != 0) continue;
			apply_dv(b, b.dv, i);
		}
	}
	if (!b.asleep) return;
	__syncthreads();
	bool w = false;
	for (int r = r0 + lane; r < r1; r += 64) {
		int ii[2];
		row_bodies(C, J, rows[2 * r], rows[2 * r + 1], ii);
		for (int q = 0; q < 2; q++) w = w || (b.movement[ii[q]] == 0 && wakes(b, S.sl.wake_vel, ii[q]));
	}
	if (__syncthreads_or(w)) {
		for (int x = S.voff[t] + lane; x < S.voff[t] + S.vcnt[t]; x += 64) wake_body(b, S.sl, S.vlist[x]);
	} else {
		for (int r = r0 + lane; r < r1; r += 64) {
			int ii[2];
			row_bodies(C, J, rows[2 * r], rows[2 * r + 1], ii);
			for (int q = 0; q < 2; q++)
				if (b.movement[ii[q]] == 0) zero_asleep(b, ii[q]);
		}
	}
}

// One colour of one sweep: each thread runs the contacts of its edges in cache
// order; the colour index MAXC holds edges that did not fit a colour and runs
// on thread 0 alone.  kind 0 warm start, 1 velocity sweep, 2 position sweep.
template <class T, class D, class Make>
__device__ __forceinline__ void colour_phase(Contacts& C, const float* iinv32, const int4* desc, const uint32_t* lmask, int e0, int beg, int end,
                                             bool serial, int kind, Make make) {
	int stride = serial ? 1 : blockDim.x;
	int start = serial ? (threadIdx.x == 0 ? beg : end) : beg + threadIdx.x;
	const int off = kind == 2 ? 6 : 0;
	for (int x = start; x < end; x += stride) {
		int4 d = desc[e0 + x];
		uint32_t lm = lmask[e0 + x];
		const int cnt = d.y & 0xFFFF;
		const bool lng = (d.y & ED_LONG) != 0;
		D MA = make(d.z, (d.y & ED_DYN_A) != 0), MB = make(d.w, ((uint32_t)d.y & ED_DYN_B) != 0);
		DvReg<T> A, B;
		A.dyn = MA.dyn;
		B.dyn = MB.dyn;
		float ia[6] = {0, 0, 0, 0, 0, 0}, ib[6] = {0, 0, 0, 0, 0, 0};
		if constexpr (sizeof(T) == 4) {
			ld_iinv32(iinv32, d.z, ia);
			ld_iinv32(iinv32, d.w, ib);
		}
#pragma unroll
		for (int c = 0; c < 6; c++) {
			A.v[c] = MA.get(off + c);
			B.v[c] = MB.get(off + c);
		}
		for (int j = 0; j < cnt; j++) {
			int k = d.x + j;
			bool live = lng ? C.live[k] != 0 : ((lm >> j) & 1u) != 0;
			if (!live) continue;
			contact_phase(rec<T>(C, k), A, B, kind, ia, ib);
		}
#pragma unroll
		for (int c = 0; c < 6; c++) {
			if (A.dyn) MA.set(off + c, A.v[c]);
			if (B.dyn) MB.set(off + c, B.v[c]);
		}
	}
}

// The coloured solve of scene s by one workgroup.  Global-memory variant
// (scenes whose velocity deltas do not fit LDS): the deltas live in b.dv
// viewed as [12][n] values of T.
template <class T>
__device__ __forceinline__ void pgs_scene_global(Bodies& b, Contacts& C, int s, const int* scene_body_off, const int* edge_range,
                                                 const int4* desc, const uint32_t* lmask, const int* color_off, const int* ncolors,
                                                 const Params& P) {
	const int b0 = scene_body_off[s], b1 = scene_body_off[s + 1];
	const int N = b.n;
	T* dv = reinterpret_cast<T*>(b.dv);
	for (int i = b0 + threadIdx.x; i < b1; i += blockDim.x)
		for (int c = 0; c < 12; c++) dv[c * N + i] = 0;
	__syncthreads();
	const int e0 = edge_range[2 * s];
	const int* co = color_off + (size_t)s * (MAXC + 2);
	const int nc = ncolors[s];
	auto make = [&](int i, bool dyn) { return DvGlobal<T>{dv, i, N, dyn}; };
	int total = 1 + P.vel_iters + P.pos_iters;
	for (int ph = 0; ph < total; ph++) {
		int kind = ph == 0 ? 0 : (ph <= P.vel_iters ? 1 : 2);
		for (int c = 0; c < nc; c++) {
			colour_phase<T, DvGlobal<T>>(C, b.iinv32, desc, lmask, e0, co[c], co[c + 1], c == MAXC, kind, make);
			__syncthreads();
		}
	}
	for (int i = b0 + threadIdx.x; i < b1; i += blockDim.x)
		if (b.movement[i] == 0) apply_dv(b, dv, i);
	__syncthreads();
}

// LDS variant: 6 values of T per body, the velocity half during warm start and
// velocity sweeps, then the pseudo-velocity half during position sweeps.
template <class T>
__device__ __forceinline__ void pgs_scene_lds(Bodies& b, Contacts& C, T* sdv, int s, const int* scene_body_off, const int* edge_range,
                                              const int4* desc, const uint32_t* lmask, const int* color_off, const int* ncolors,
                                              const Params& P) {
	const int b0 = scene_body_off[s], nbs = scene_body_off[s + 1] - b0;
	const int N = b.n;
	T* dv = reinterpret_cast<T*>(b.dv);
	const int e0 = edge_range[2 * s];
	const int* co = color_off + (size_t)s * (MAXC + 2);
	const int nc = ncolors[s];
	auto make = [&](int i, bool dyn) { return DvLds<T>{sdv, i - b0, nbs, dyn}; };
	for (int half = 0; half < 2; half++) {
		for (int x = threadIdx.x; x < 6 * nbs; x += blockDim.x) sdv[x] = 0;
		__syncthreads();
		int ph0 = half == 0 ? 0 : 1 + P.vel_iters;
		int ph1 = half == 0 ? 1 + P.vel_iters : 1 + P.vel_iters + P.pos_iters;
		for (int ph = ph0; ph < ph1; ph++) {
			int kind = ph == 0 ? 0 : (ph <= P.vel_iters ? 1 : 2);
			for (int c = 0; c < nc; c++) {
				colour_phase<T, DvLds<T>>(C, b.iinv32, desc, lmask, e0, co[c], co[c + 1], c == MAXC, kind, make);
				__syncthreads();
			}
		}
		// write the half back (the residual reads it) and apply it (PhysicsEngine.cpp:1113-1136)
		for (int li = threadIdx.x; li < nbs; li += blockDim.x) {
			int i = b0 + li;
			for (int c = 0; c < 6; c++) dv[(6 * half + c) * N + i] = sdv[c * nbs + li];
			if (b.movement[i] != 0) continue;
			for (int c = 0; c < 3; c++) {
				if (half == 0) {
					b.vel[c * N + i] += (double)sdv[c * nbs + li];
					b.angv[c * N + i] += (double)sdv[(3 + c) * nbs + li];
				} else {
					b.pvel[c * N + i] += (double)sdv[c * nbs + li];
					b.pang[c * N + i] += (double)sdv[(3 + c) * nbs + li];
				}
			}
		}
		__syncthreads();
	}
}

// One coloured solve, one workgroup per scene (scene_list, or every scene).
template <class T, int MAXT>
__global__ __launch_bounds__(MAXT) void k_pgs_colored(Bodies b, Contacts C, const int* scene_body_off, const int* edge_range, const int4* desc,
                                                     const uint32_t* lmask, const int* color_off, const int* ncolors, Params P,
                                                     const int* scene_list) {
	const int s = scene_list ? scene_list[blockIdx.x] : blockIdx.x;
	pgs_scene_global<T>(b, C, s, scene_body_off, edge_range, desc, lmask, color_off, ncolors, P);
}
template <class T, int MAXT>
__global__ __launch_bounds__(MAXT) void k_pgs_colored_lds(Bodies b, Contacts C, const int* scene_list, const int* scene_body_off, const int* edge_range,
                                                         const int4* desc, const uint32_t* lmask, const int* color_off, const int* ncolors,
                                                         Params P) {
	extern __shared__ double lds_raw[];
	pgs_scene_lds<T>(b, C, reinterpret_cast<T*>(lds_raw), scene_list[blockIdx.x], scene_body_off, edge_range, desc, lmask, color_off,
	                 ncolors, P);
}

template <class T>
__device__ __forceinline__ void writeback_contact(Contacts& C, int k) {
	const T* R = rec<T>(C, k);
	C.imp_n[k] = (double)R[CP_R_IMP];
	C.imp_t[k] = (double)R[CP_R_FIMP];
	C.imp_t[C.cap + k] = (double)R[CP_R_FIMP + 1];
	C.sready[k] = R[CP_R_SREADY] != T(0);
}

// The whole substep loop of one scene in one workgroup (P.substeps times:
// rows of the scene's live contacts, the coloured solve, the impulses back
// into the contact cache, integration of the scene's dynamic bodies).  A
// scene's contacts are one contiguous range of the cache (its edges'), and a
// workgroup touches only its own scene's bodies, contacts and rows, so the
// substeps of different scenes need no synchronisation between them; the
// arithmetic is the per-substep kernels', in the same order.
template <class T, bool LDS>
__device__ __forceinline__ void substeps_scene(Bodies& b, Contacts& C, T* sdv, int s, const int* scene_body_off, const int* edge_range,
                                               const int* efirst, const int* ecnt, const int4* desc, const uint32_t* lmask,
                                               const int* color_off, const int* ncolors, const Params& P) {
	const int b0 = scene_body_off[s], b1 = scene_body_off[s + 1];
	const int e0 = edge_range[2 * s], e1 = edge_range[2 * s + 1];
	const int c0 = e1 > e0 ? efirst[e0] : 0, c1 = e1 > e0 ? efirst[e1 - 1] + ecnt[e1 - 1] : 0;
	for (int sub = 0; sub < P.substeps; sub++) {
		if constexpr (sizeof(T) == 4) {
			for (int i = b0 + threadIdx.x; i < b1; i += blockDim.x) iinv32_body(b, i);
			__syncthreads();
		}
		for (int k = c0 + threadIdx.x; k < c1; k += blockDim.x)
			if (C.live[k]) build_row<T>(b, C, k, P);
		__syncthreads();
		if (LDS) pgs_scene_lds<T>(b, C, sdv, s, scene_body_off, edge_range, desc, lmask, color_off, ncolors, P);
		else pgs_scene_global<T>(b, C, s, scene_body_off, edge_range, desc, lmask, color_off, ncolors, P);
		for (int k = c0 + threadIdx.x; k < c1; k += blockDim.x)
			if (C.live[k]) writeback_contact<T>(C, k);
		for (int i = b0 + threadIdx.x; i < b1; i += blockDim.x)
			if (b.movement[i] == 0) integrate_body(b, P, i);
		__syncthreads();
	}
}
template <class T, int MAXT>
__global__ __launch_bounds__(MAXT) void k_substeps_lds(Bodies b, Contacts C, const int* scene_list, const int* scene_body_off,
                                                      const int* edge_range, const int* efirst, const int* ecnt, const int4* desc,
                                                      const uint32_t* lmask, const int* color_off, const int* ncolors, Params P) {
	extern __shared__ double lds_raw[];
	substeps_scene<T, true>(b, C, reinterpret_cast<T*>(lds_raw), scene_list[blockIdx.x], scene_body_off, edge_range, efirst, ecnt, desc,
	                        lmask, color_off, ncolors, P);
}
template <class T, int MAXT>
__global__ __launch_bounds__(MAXT) void k_substeps_global(Bodies b, Contacts C, const int* scene_list, const int* scene_body_off,
                                                         const int* edge_range, const int* efirst, const int* ecnt, const int4* desc,
                                                         const uint32_t* lmask, const int* color_off, const int* ncolors, Params P) {
	substeps_scene<T, false>(b, C, nullptr, scene_list[blockIdx.x], scene_body_off, edge_range, efirst, ecnt, desc, lmask, color_off,
	                         ncolors, P);
}

// Accumulated impulses and the static flag back into the persistent FP64
// contact cache (warm start of the next substep; PersistentConstraint.cpp:201-215).
template <class T>
__global__ void k_rows_writeback(Contacts C, int nc, const int* ncp) {
	const int n = ncp ? *ncp : nc;
	for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x)
		if (C.live[k]) writeback_contact<T>(C, k);
}

// FP64 natural-map residual of the velocity rows after the sweeps, evaluated
// in FP64 from the stored rows and deltas: acc[0] += sum r^2, acc[1] += sum
// lambda^2, acc[2] += rows.  Runs right after the PGS launch (reads b.dv).
template <class T>
__global__ void k_pgs_residual(Bodies b, Contacts C, int nc, double* acc) {
	int k = blockIdx.x * blockDim.x + threadIdx.x;
	double r2 = 0, l2 = 0, rows = 0;
	if (k < nc && C.live[k]) {
		const int N = b.n;
		const T* R = rec<T>(C, k);
		const T* dv = reinterpret_cast<const T*>(b.dv);
		int i1 = (int)(C.key[k] >> 32), i2 = (int)(C.key[k] & 0xFFFFFFFFu);
		RowRegs<T> g;
		if constexpr (sizeof(T) == 4) {  // the Jacobians do not involve the inertia
			float r[CP_RSF];
			const float z[6] = {0, 0, 0, 0, 0, 0};
			ldv(R, r);
			compact_normal(r, z, z, g);
			compact_friction(r, z, z, g);
		} else {
			T bs[12];
			ldv(R + CP_R_BASIS, bs);
			load_normal(R, bs, g);
			load_friction(R, bs, g);
		}
		double x = 0, y = 0;
		for (int q = 0; q < 6; q++) {
			x += (double)dv[q * N + i1] * (double)g.j[q];
			y += (double)dv[q * N + i2] * (double)g.j[6 + q];
		}
		double ln = R[CP_R_IMP];
		double z = ln + (double)R[CP_R_INV] * ((double)R[CP_R_TGT] - (x + y));
		double rn = ln - (z < 0 ? 0 : z);
		double cur[2];
		for (int r = 0; r < 2; r++) {
			double a = 0, c = 0;
			for (int q = 0; q < 6; q++) {
				a += (double)dv[q * N + i1] * (double)g.fj[r * 6 + q];
				c += (double)dv[q * N + i2] * (double)g.fj[12 + r * 6 + q];
			}
			cur[r] = a + c;
		}
		double d0 = (double)R[CP_R_FTGT] - cur[0], d1 = (double)R[CP_R_FTGT + 1] - cur[1];
		double t0 = R[CP_R_FIMP], t1 = R[CP_R_FIMP + 1];
		double z0 = t0 + (double)R[CP_R_FINV] * d0 + (double)R[CP_R_FINV + 1] * d1;
		double z1 = t1 + (double)R[CP_R_FINV + 2] * d0 + (double)R[CP_R_FINV + 3] * d1;
		double lim = (double)R[CP_R_MU] * (ln > 0 ? ln : 0);
		double m = sqrt(z0 * z0 + z1 * z1);
		if (m > lim && m > 0) { z0 *= lim / m; z1 *= lim / m; }
		double q0 = t0 - z0, q1 = t1 - z1;
		r2 = rn * rn + q0 * q0 + q1 * q1;
		l2 = ln * ln + t0 * t0 + t1 * t1;
		rows = 3;
	}
	for (int off = 32; off >= 1; off >>= 1) {  // wave reduction, one atomic per wave
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

// With ncp, nc is an upper bound on the device count (the grid is sized for
// it, capped; threads stride over the rest).
static inline int grid_for(int n, int t) { return std::max(1, std::min(nblk(n, t), 4096)); }
void launch_build_rows(const Bodies& b, const Contacts& C, int nc, const Params& P, hipStream_t s, const int* ncp) {
	if (!nc) return;
	int g = xcd_grid(ncp ? grid_for(nc, 64) : nblk(nc, 64));
	if (P.precision == CP_PREC_SPLIT) hipLaunchKernelGGL(k_build_rows<float>, dim3(g), dim3(64), 0, s, b, C, nc, ncp, P);
	else hipLaunchKernelGGL(k_build_rows<double>, dim3(g), dim3(64), 0, s, b, C, nc, ncp, P);
}
void launch_iinv32(const Bodies& b, hipStream_t s) {
	if (b.nl) hipLaunchKernelGGL(k_iinv32, dim3(nblk(b.nl, 256)), dim3(256), 0, s, b);
}
void launch_pgs_islands(const Bodies& b, const Contacts& C, const Joints& J, const HSysDev& H, const IslandSched& S, const Params& P,
                        int* applied, double* lin, int first, bool colored, hipStream_t s) {
	if (!S.n) return;
	if (colored) {
		static bool attr = false;
		if (!attr) {
			(void)hipFuncSetAttribute((const void*)k_pgs_islands, hipFuncAttributeMaxDynamicSharedMemorySize, 64 * 1024);
			attr = true;
		}
		size_t lds = S.cloc ? (size_t)12 * S.lds_bodies * sizeof(double) : 0;
		hipLaunchKernelGGL(k_pgs_islands, dim3(S.n), dim3(64), lds, s, b, C, J, H, S, P, applied, lin, first);
	}
	else hipLaunchKernelGGL(k_pgs_ordered, dim3(nblk(S.n, 64)), dim3(64), 0, s, b, C, J, H, S, P, applied, lin, first);
}
// Workgroups of more than 256 threads (up to 1024: a scene's colour then runs
// on 16 waves) use the kernels instantiated with a wider launch bound (512:
// the FP64 kernels keep their registers; 1024: 4 waves per SIMD).
template <class T>
static void pgs_colored(const Bodies& b, const Contacts& C, const int* scene_body_off, int nb, const int* edge_range, const int4* desc,
                        const uint32_t* lmask, const int* color_off, const int* ncolors, const Params& P, int threads, hipStream_t s,
                        const int* scene_list) {
	if (threads > 512)
		hipLaunchKernelGGL((k_pgs_colored<T, 1024>), dim3(nb), dim3(threads), 0, s, b, C, scene_body_off, edge_range, desc, lmask, color_off,
		                   ncolors, P, scene_list);
	else if (threads > 256)
		hipLaunchKernelGGL((k_pgs_colored<T, 512>), dim3(nb), dim3(threads), 0, s, b, C, scene_body_off, edge_range, desc, lmask, color_off,
		                   ncolors, P, scene_list);
	else
		hipLaunchKernelGGL((k_pgs_colored<T, 256>), dim3(nb), dim3(threads), 0, s, b, C, scene_body_off, edge_range, desc, lmask, color_off,
		                   ncolors, P, scene_list);
}
void launch_pgs_colored(const Bodies& b, const Contacts& C, const int* scene_body_off, int nscenes, const int* edge_range, const int4* desc,
                        const uint32_t* lmask, const int* color_off, const int* ncolors, const Params& P, int threads, hipStream_t s,
                        const int* scene_list, int nlist) {
	int nb = scene_list ? nlist : nscenes;
	if (!nb) return;
	if (P.precision == CP_PREC_SPLIT)
		pgs_colored<float>(b, C, scene_body_off, nb, edge_range, desc, lmask, color_off, ncolors, P, threads, s, scene_list);
	else
		pgs_colored<double>(b, C, scene_body_off, nb, edge_range, desc, lmask, color_off, ncolors, P, threads, s, scene_list);
}
template <class T, int MAXT>
static void pgs_colored_lds(const Bodies& b, const Contacts& C, const int* scene_list, int nlist, int max_bodies, const int* scene_body_off,
                            const int* edge_range, const int4* desc, const uint32_t* lmask, const int* color_off, const int* ncolors,
                            const Params& P, int threads, hipStream_t s) {
	static bool attr = false;
	if (!attr) {
		(void)hipFuncSetAttribute((const void*)k_pgs_colored_lds<T, MAXT>, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
		attr = true;
	}
	hipLaunchKernelGGL((k_pgs_colored_lds<T, MAXT>), dim3(nlist), dim3(threads), (size_t)6 * max_bodies * sizeof(T), s, b, C, scene_list,
	                   scene_body_off, edge_range, desc, lmask, color_off, ncolors, P);
}
void launch_pgs_colored_lds(const Bodies& b, const Contacts& C, const int* scene_list, int nlist, int max_bodies, const int* scene_body_off,
                            const int* edge_range, const int4* desc, const uint32_t* lmask, const int* color_off, const int* ncolors,
                            const Params& P, int threads, hipStream_t s) {
	if (!nlist) return;
#define CP_LDS_ARGS b, C, scene_list, nlist, max_bodies, scene_body_off, edge_range, desc, lmask, color_off, ncolors, P, threads, s
	if (P.precision == CP_PREC_SPLIT) {
		if (threads > 512) pgs_colored_lds<float, 1024>(CP_LDS_ARGS);
		else if (threads > 256) pgs_colored_lds<float, 512>(CP_LDS_ARGS);
		else pgs_colored_lds<float, 256>(CP_LDS_ARGS);
	} else {
		if (threads > 512) pgs_colored_lds<double, 1024>(CP_LDS_ARGS);
		else if (threads > 256) pgs_colored_lds<double, 512>(CP_LDS_ARGS);
		else pgs_colored_lds<double, 256>(CP_LDS_ARGS);
	}
#undef CP_LDS_ARGS
}
template <class T, int MAXT>
static void substeps_lds(const Bodies& b, const Contacts& C, const int* scene_list, int nlist, int max_bodies, const int* scene_body_off,
                         const int* edge_range, const int* efirst, const int* ecnt, const int4* desc, const uint32_t* lmask,
                         const int* color_off, const int* ncolors, const Params& P, int threads, hipStream_t s) {
	static bool attr = false;
	if (!attr) {
		(void)hipFuncSetAttribute((const void*)k_substeps_lds<T, MAXT>, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
		attr = true;
	}
	hipLaunchKernelGGL((k_substeps_lds<T, MAXT>), dim3(nlist), dim3(threads), (size_t)6 * max_bodies * sizeof(T), s, b, C, scene_list,
	                   scene_body_off, edge_range, efirst, ecnt, desc, lmask, color_off, ncolors, P);
}
void launch_substeps_lds(const Bodies& b, const Contacts& C, const int* scene_list, int nlist, int max_bodies, const int* scene_body_off,
                         const int* edge_range, const int* efirst, const int* ecnt, const int4* desc, const uint32_t* lmask,
                         const int* color_off, const int* ncolors, const Params& P, int threads, hipStream_t s) {
	if (!nlist) return;
#define CP_SUB_ARGS b, C, scene_list, nlist, max_bodies, scene_body_off, edge_range, efirst, ecnt, desc, lmask, color_off, ncolors, P, threads, s
	if (P.precision == CP_PREC_SPLIT) {
		if (threads > 512) substeps_lds<float, 1024>(CP_SUB_ARGS);
		else if (threads > 256) substeps_lds<float, 512>(CP_SUB_ARGS);
		else substeps_lds<float, 256>(CP_SUB_ARGS);
	} else {
		if (threads > 512) substeps_lds<double, 1024>(CP_SUB_ARGS);
		else if (threads > 256) substeps_lds<double, 512>(CP_SUB_ARGS);
		else substeps_lds<double, 256>(CP_SUB_ARGS);
	}
#undef CP_SUB_ARGS
}
void launch_substeps_global(const Bodies& b, const Contacts& C, const int* scene_list, int nlist, const int* scene_body_off, const int* edge_range,
                            const int* efirst, const int* ecnt, const int4* desc, const uint32_t* lmask, const int* color_off,
                            const int* ncolors, const Params& P, int threads, hipStream_t s) {
	if (!nlist) return;
#define CP_SUB_ARGS dim3(nlist), dim3(threads), 0, s, b, C, scene_list, scene_body_off, edge_range, efirst, ecnt, desc, lmask, color_off, ncolors, P
	if (P.precision == CP_PREC_SPLIT) {
		if (threads > 512) hipLaunchKernelGGL((k_substeps_global<float, 1024>), CP_SUB_ARGS);
		else if (threads > 256) hipLaunchKernelGGL((k_substeps_global<float, 512>), CP_SUB_ARGS);
		else hipLaunchKernelGGL((k_substeps_global<float, 256>), CP_SUB_ARGS);
	} else {
		if (threads > 512) hipLaunchKernelGGL((k_substeps_global<double, 1024>), CP_SUB_ARGS);
		else if (threads > 256) hipLaunchKernelGGL((k_substeps_global<double, 512>), CP_SUB_ARGS);
		else hipLaunchKernelGGL((k_substeps_global<double, 256>), CP_SUB_ARGS);
	}
#undef CP_SUB_ARGS
}
void launch_rows_writeback(const Contacts& C, int nc, const Params& P, hipStream_t s, const int* ncp) {
	if (!nc) return;
	int g = ncp ? grid_for(nc, 256) : nblk(nc, 256);
	if (P.precision == CP_PREC_SPLIT) hipLaunchKernelGGL(k_rows_writeback<float>, dim3(g), dim3(256), 0, s, C, nc, ncp);
	else hipLaunchKernelGGL(k_rows_writeback<double>, dim3(g), dim3(256), 0, s, C, nc, ncp);
}
void launch_pgs_residual(const Bodies& b, const Contacts& C, int nc, const Params& P, double* acc, hipStream_t s) {
	if (!nc) return;
	if (P.precision == CP_PREC_SPLIT) hipLaunchKernelGGL(k_pgs_residual<float>, dim3(nblk(nc, 256)), dim3(256), 0, s, b, C, nc, acc);
	else hipLaunchKernelGGL(k_pgs_residual<double>, dim3(nblk(nc, 256)), dim3(256), 0, s, b, C, nc, acc);
}

}  // namespace cpx
