// Joints (ball-socket, hinge) and the holonomic block solver on the device.
//
// Rows: BallSocket / Hinge::updateSolverConstraints (PersistentConstraint.cpp:
// 225-231, :274-285) -> BallSocketConstraint / HingeConstraint (ConstraintSolver.cpp:
// 382-405, :468-550); degree n = 3 or 5, NMat accumulation order (NMat.h), so
// FP64 rows match the reference bit for bit.  One thread per joint.
//
// Holonomic systems (HolonomicBlockSolver.cpp): the symbolic part (elimination
// order, block table) is host work done when the joint topology changes; the
// numeric block LDL factor (:459-543) runs one wave per system on its blocks
// staged in LDS (one thread per system for systems too large for LDS), and the solve
// d = L^-T D^-1 L^-1 (t - c) (:223-355) runs inside the island's PGS thread
// after the sweeps of the first substep.
//
// Roofline: per joint per substep ~600 B of body state in, 160 doubles of rows
// out; per system the factor touches sum(d_i^2 + d_i d_(i+1)) doubles (5-7 KB
// per 16-link chain).  Latency bound at these sizes.
#include "k_hsys.h"

namespace cpx {

__device__ __forceinline__ V3 jld3(const double* a, int i, int n) { return V3{a[i], a[n + i], a[2 * n + i]}; }
__device__ __forceinline__ void jst3(double* a, int i, int n, V3 v) {
	a[i] = v.x;
	a[n + i] = v.y;
	a[2 * n + i] = v.z;
}

__device__ __forceinline__ double pos_coeff(double s, double dt) {  // posCorrectCoeff (PersistentConstraint.cpp:199)
	double a = s * dt, b = 1.0 / dt;
	return (b < a) ? b : a;
}

// rowMajorOrderInverse: 1x1, the 3x3 closed form (NMat.cpp:24-36), otherwise
// unpivoted Gauss-Jordan (NMat.h:121-147).
__device__ void inverse_n(int n, const double* src, double* dst) {
	if (n == 1) {
		dst[0] = 1.0 / src[0];
		return;
	}
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
	for (int k = 0; k < n * n; k++) {
		cp[k] = src[k];
		t[k] = 0;
	}
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

// x S(v) for a row vector x, S the cross-product matrix crossMat (NMat.cpp:
// 62-70), accumulated as NMat's product accumulates (NMat.h:95-103).
__device__ __forceinline__ void row_cross(const double* x, V3 v, double* out) {
	const double S[3][3] = {{0, -v.z, v.y}, {v.z, 0, -v.x}, {-v.y, v.x, 0}};
	for (int c = 0; c < 3; c++) {
		double acc = 0;
		for (int q = 0; q < 3; q++) acc += x[q] * S[q][c];
		out[c] = acc;
	}
}
__device__ __forceinline__ void put3(double (*m)[6], int r, int c0, const double* x) {
	for (int c = 0; c < 3; c++) m[r][c0 + c] = x[c];
}
// I (idenMat<3>) or -I into rows r0.., columns c0..; -I keeps its negative zeros.
__device__ __forceinline__ void put_iden(double (*m)[6], int r0, int c0, bool negative) {
	for (int r = 0; r < 3; r++)
		for (int c = 0; c < 3; c++) {
			double e = (r == c) ? 1.0 : 0.0;
			m[r0 + r][c0 + c] = negative ? -e : e;
		}
}
// crossMat(v) or -crossMat(v) into rows r0.., columns c0...
__device__ __forceinline__ void put_cross(double (*m)[6], int r0, int c0, V3 v, bool negative) {
	const double S[3][3] = {{0, -v.z, v.y}, {v.z, 0, -v.x}, {-v.y, v.x, 0}};
	for (int r = 0; r < 3; r++)
		for (int c = 0; c < 3; c++) m[r0 + r][c0 + c] = negative ? -S[r][c] : S[r][c];
}
// Quaternion::angleTo (Quaternion.h:35-39) and getRotAxis (:41-49): the
// rotation correction of the weld and slider rows (ConstraintSolver.cpp:
// 688-696, :888-896), axis * angle scaled by h.
__device__ V3 rot_correct(Q a, Q b, double h) {
	Q diff = q_mul(b, q_conj(a));
	double dot = fabs(a.r * b.r + a.i * b.i + a.j * b.j + a.k * b.k);
	double angle = dot > 0.99999999999 ? 0 : 2 * acos(fabs(a.r * b.r + a.i * b.i + a.j * b.j + a.k * b.k));
	V3 axis{0, 0, 0};
	if (angle != 0) {
		V3 ax{diff.i, diff.j, diff.k};
		V3 nx = normalize(ax);
		axis = (diff.r < 0 ? V3{-nx.x, -nx.y, -nx.z} : nx) * h;
	}
	return V3{axis.x * angle, axis.y * angle, axis.z * angle};
}

// Taylor approximation of 1 - cos f, the limits' position error
// (ConstraintSolver.cpp:425-428, :455-457).
__device__ __forceinline__ double limit_error(double f) {
	return -f * f * f * f * f * f * f * f / 40320.0 + f * f * f * f * f * f / 720.0 - f * f * f * f / 24.0 + f * f / 2.0;
}
#define CP_QUAT_PI 3.1415926535  // PI of Quaternion.h:4, the twist angle's unwrap

// Motor::calculatePosition (PersistentConstraint.cpp:30-71): the angle
// predicted from the relative angular velocity, snapped to the measured angle
// of b2's comparison axis in b1's reference plane (the candidate nearest the
// prediction among the measured angle shifted by -2PI, 0, +2PI).
__device__ double motor_position(const Joints& J, int k, Q qa, Q qb, V3 axis, V3 av1, V3 av2, double dt, double prev) {
	const int K = J.n;
	V3 u1, w1, u2, tmp;
	perp_basis(jld3(J.ax1, k, K), &u1, &w1);  // Motor ctor (:8-13)
	perp_basis(jld3(J.ax2, k, K), &u2, &tmp);
	V3 ru = q_rotate(qa, u1), rw = q_rotate(qa, w1), cmp = q_rotate(qb, u2);
	double rel = dot(axis, av1) - dot(axis, av2);
	double pred = prev + rel * dt;
	if (fabs(dot(cmp, axis)) > 0.999) return pred;
	V3 cn = normalize(cmp - axis * dot(axis, cmp));
	double dv = dot(cn, ru);
	if (dv > 1.0) dv = 1.0;
	if (dv < -1.0) dv = -1.0;
	double relang = 2 * CP_QUAT_PI - acos(dv);
	if (dot(cn, rw) < 0) relang = 2 * CP_QUAT_PI - relang;
	double r = fmod(pred, 2 * CP_QUAT_PI);
	double n2 = pred - r;
	const double cand[3] = {n2 + relang - 2 * CP_QUAT_PI, n2 + relang, n2 + relang + 2 * CP_QUAT_PI};
	double closest = __builtin_huge_val();
	for (int c = 0; c < 3; c++)
		if (fabs(pred - cand[c]) < fabs(pred - closest)) closest = cand[c];
	return closest;
}

// getConstraintTargetVelocityValue of the motor (:77-108) and the piston
// (:153-185).  vrel: the relative velocity along the axis as each measures
// it; inv_sum: inverse inertia (motor) or inverse mass (piston) along it;
// the reference tests abs(impulse < EPS), i.e. impulse < EPS.
__device__ double drive_target(const Joints& J, int k, double vrel, double inv_sum, double imp, double pos, double prev_vel) {
	const int K = J.n;
	const int mode = (int)J.drive[k];
	const double mx = J.drive[K + k];
	if (mode == 0) return 0;
	if (mode == 1) return mx > 0 ? __builtin_huge_val() : -__builtin_huge_val();
	if (mode == 2) return J.drive[2 * K + k];
	double diff = pos - J.drive[3 * K + k];
	double dir = (diff > 0) ? -1 : 1;
	double vd = fabs(vrel - prev_vel);
	const double EPS = 0.0000001;
	double inertia = (vd < EPS || imp < EPS) ? 1.0 / inv_sum : fabs(imp) / vd;
	return dir * sqrt(2 * fabs(diff) * mx / inertia);
}

// Solver rows of joint k from the current body state: the joint's
// updateSolverConstraints (PersistentConstraint.cpp:217-354, warm start
// coefficient 1) and its constraint's constructor (ConstraintSolver.cpp:
// 348-900).  reproject: hinges, sliders and sliding hinges re-project their
// stored impulses onto the new basis (ConstraintSolver.cpp:475-476, :617-618,
// :787-788); false for holonomic rows rebuilt from a stored state.
__device__ void build_joint(const Bodies& b, Joints& J, int k, double sub_dt, double cfm, bool reproject) {
	const int N = b.n, K = J.n;
	const int n = J.type[k];
	const int ia = J.a[k], ib = J.b[k];
	double* R = J.row + (size_t)k * CP_JRS;
	Q qa{b.q[ia], b.q[N + ia], b.q[2 * N + ia], b.q[3 * N + ia]};
	Q qb{b.q[ib], b.q[N + ib], b.q[2 * N + ib], b.q[3 * N + ib]};
	V3 ca = jld3(b.com, ia, N), cb = jld3(b.com, ib, N);
	V3 pa = ca + q_rotate(qa, jld3(J.tp1, k, K));  // getTrackedP (RigidBody.cpp:139-144)
	V3 pb = cb + q_rotate(qb, jld3(J.tp2, k, K));
	bool da = b.movement[ia] == 0, db = b.movement[ib] == 0;
	double Ma[6][6], Mb[6][6];
	{
		double ima = da ? 1.0 / b.mass[ia] : 0, imb = db ? 1.0 / b.mass[ib] : 0;
		for (int r = 0; r < 6; r++) for (int c = 0; c < 6; c++) { Ma[r][c] = 0; Mb[r][c] = 0; }
		for (int r = 0; r < 3; r++)
			for (int c = 0; c < 3; c++) {
				Ma[r][c] = (r == c ? 1.0 : 0.0) * ima;
				Mb[r][c] = (r == c ? 1.0 : 0.0) * imb;
				Ma[3 + r][3 + c] = da ? b.Iinv[(3 * r + c) * N + ia] : 0.0;
				Mb[3 + r][3 + c] = db ? b.Iinv[(3 * r + c) * N + ib] : 0.0;
			}
	}
	double start[CP_JMAX];
	for (int c = 0; c < n; c++) start[c] = J.imp[c * K + k] * 1.0;
	const int kind = J.kind[k];
	// an inequality row: linear and angular parts for each body, pseudo target
	V3 il_a{0, 0, 0}, lim_a{0, 0, 0}, il_b{0, 0, 0}, lim_b{0, 0, 0};
	double ipt = 0;
	auto vel_of = [&](const double* v, int i) { return b.movement[i] == 1 ? V3{0, 0, 0} : jld3(v, i, N); };  // getVel / getAngVel
	if (kind == 8) {
		// ConeRotationLimit::updateSolverConstraints (PersistentConstraint.cpp:233-247):
		// inside the cone the limit clears its impulse and keeps its rows
		V3 d1 = q_rotate(qa, jld3(J.ax1, k, K)), d2 = q_rotate(qb, jld3(J.ax2, k, K));
		double angle = acos(dot(d1, d2));
		const double lim = J.tp1[k];
		bool inactive = angle < lim;
		J.u[k] = inactive ? 1.0 : 0.0;
		if (inactive) {
			J.imp[k] = 0.0;
			return;
		}
		// ConeLimitConstraint (ConstraintSolver.cpp:410-430)
		V3 h = normalize(cross(d1, d2));
		lim_a = h;
		lim_b = V3{-h.x, -h.y, -h.z};
		ipt = limit_error(angle - lim) * J.rs[k];  // the raw rotation hardness, not posCorrectCoeff
		J.proj[k] = CP_PROJ_LOWER;
	} else if (kind == 9) {
		// TwistLimit::updateSolverConstraints (PersistentConstraint.cpp:249-272);
		// the angle as getCurrentAngle (:115-132) measures it: the reference
		// rotates b1's local w axis by b2's orientation, and the b2-axis
		// alignment it computes is not used
		M3 r1 = q_rot(qa), r2 = q_rot(qb);
		V3 axl = jld3(J.ax1, k, K), wl, ul;
		perp_basis(axl, &wl, &ul);  // the constructor's getPerpendicularBasis(&w, &u)
		V3 t1 = r1 * axl, u1 = r1 * ul, w1 = r1 * wl;
		V3 t2 = r2 * jld3(J.ax2, k, K);
		V3 r2w = r2 * wl;
		double prev = J.u[K + k];
		double cur = atan2(dot(r2w, u1), dot(r2w, w1));
		while (prev - cur > CP_QUAT_PI) cur += 2 * CP_QUAT_PI;
		while (prev - cur < -CP_QUAT_PI) cur -= 2 * CP_QUAT_PI;
		J.u[K + k] = cur;
		const double mn = J.tp1[k], mx = J.tp1[K + k];
		bool inactive = mn <= cur && cur <= mx;
		J.u[k] = inactive ? 1.0 : 0.0;
		if (inactive) {
			J.imp[k] = 0.0;
			return;
		}
		// TwistLimitConstraint (ConstraintSolver.cpp:439-459)
		bool above = cur > mx;
		lim_a = V3{-t1.x, -t1.y, -t1.z};
		lim_b = t2;
		ipt = limit_error(above ? cur - mx : mn - cur) * J.rs[k];
		J.proj[k] = above ? CP_PROJ_UPPER : CP_PROJ_LOWER;
	} else if (kind == 10) {
		// Motor::updateSolverConstraints (PersistentConstraint.cpp:320-329) ->
		// MotorConstraint (ConstraintSolver.cpp:555-584); prev_velocity is
		// written after the build either way
		V3 axis = q_rotate(qa, jld3(J.ax1, k, K));
		V3 av1 = vel_of(b.angv, ia), av2 = vel_of(b.angv, ib);
		const double pos = motor_position(J, k, qa, qb, axis, av1, av2, sub_dt, J.u[k]);
		J.u[k] = pos;
		const double pv = dot(axis, av1) - dot(axis, av2);
		JointAct act{K, J.kind, J.u, J.drive, J.tp1};
		if (!act(k)) {
			J.u[K + k] = pv;
			return;
		}
		M3 ia3 = m3_zero(), ib3 = m3_zero();
		for (int r = 0; r < 3; r++)
			for (int c = 0; c < 3; c++) {
				if (da) ia3.m[r][c] = b.Iinv[(3 * r + c) * N + ia];
				if (db) ib3.m[r][c] = b.Iinv[(3 * r + c) * N + ib];
			}
		const double inv_sum = dot(axis, ia3 * axis) + dot(axis, ib3 * axis);
		const double target = 1.0 * drive_target(J, k, pv, inv_sum, start[0], pos, J.u[K + k]);
		const double rcc0 = pos_coeff(J.rs[k], sub_dt);
		const double mn = J.tp1[k], mx = J.tp1[K + k];
		R[CP_J_BOUND] = fabs(J.drive[K + k] * sub_dt);
		if (pos > mx) {
			ipt = (mx - pos) * rcc0;
			R[CP_J_TV] = target < 0 ? target : 0.0;  // std::min<double>(0, target)
			J.proj[k] = CP_PROJ_UPPER;
		} else if (pos < mn) {
			ipt = (mn - pos) * rcc0;
			R[CP_J_TV] = 0.0 < target ? target : 0.0;  // std::max<double>(0, target)
			J.proj[k] = CP_PROJ_LOWER;
		} else {
			ipt = 0;
			R[CP_J_TV] = target;
			J.proj[k] = CP_PROJ_CLAMP | CP_PROJ_NOPOS;
		}
		lim_a = axis;
		lim_b = V3{-axis.x, -axis.y, -axis.z};
		J.u[K + k] = pv;
	} else if (kind == 11 || kind == 12) {
		// Piston::updateSolverConstraints (PersistentConstraint.cpp:300-318)
		V3 axis = q_rotate(qa, jld3(J.ax1, k, K));
		const double pos = dot(pb, axis) - dot(pa, axis);  // Piston::calculatePosition (:149-151)
		if (kind == 12) {
			J.u[K + k] = pos;
			const double mn = J.ax2[k], mx = J.ax2[K + k];
			bool inactive = !(pos < mn || pos > mx);  // slideLimitExceeded (:191-193)
			J.u[k] = inactive ? 1.0 : 0.0;
			if (inactive) return;
			// SlideLimitConstraint (ConstraintSolver.cpp:736-766)
			const double pcc0 = pos_coeff(J.ps[k], sub_dt);
			if (pos > mx) {
				ipt = (mx - pos) * pcc0;
				J.proj[k] = CP_PROJ_UPPER;
			} else {
				ipt = (mn - pos) * pcc0;
				J.proj[k] = CP_PROJ_LOWER;
			}
		} else {
			J.u[k] = pos;
			JointAct act{K, J.kind, J.u, J.drive, J.tp1};
			if (!act(k)) return;
			// PistonConstraint (ConstraintSolver.cpp:701-721); the piston never
			// writes prev_velocity
			const double inv_sum = (da ? 1.0 / b.mass[ia] : 0.0) + (db ? 1.0 / b.mass[ib] : 0.0);
			V3 v1 = vel_of(b.vel, ia), v2 = vel_of(b.vel, ib);
			R[CP_J_TV] = drive_target(J, k, dot(axis, v1) - dot(axis, v2), inv_sum, start[0], pos, J.u[K + k]);
			R[CP_J_BOUND] = J.drive[K + k] * sub_dt;
			J.proj[k] = CP_PROJ_CLAMP | CP_PROJ_NOPOS;
		}
		il_a = V3{-axis.x, -axis.y, -axis.z};
		const double ns[3] = {-axis.x, -axis.y, -axis.z};
		double t[3];
		row_cross(ns, axis, t);
		lim_a = V3{t[0], t[1], t[2]};
		il_b = axis;
	}
	for (int c = 0; c < n; c++) J.imp[c * K + k] = start[c];
	double ja[CP_JMAX][6], jb[CP_JMAX][6];
	for (int r = 0; r < CP_JMAX; r++)
		for (int c = 0; c < 6; c++) {
			ja[r][c] = 0;
			jb[r][c] = 0;
		}
	const double pcc = pos_coeff(J.ps[k], sub_dt), rcc = pos_coeff(J.rs[k], sub_dt);
	double* pt = R + CP_J_PTGT;
	// new basis (u, w) of a basis-carrying joint and its re-projection of
	// impulses i0, i0 + 1
	auto basis = [&](V3 axis, int i0, V3* u, V3* w) {
		V3 su = jld3(J.u, k, K), sw = jld3(J.w, k, K);
		perp_basis(axis, u, w);
		if (reproject) {
			J.imp[i0 * K + k] = dot(*u, su) * start[i0] + dot(*u, sw) * start[i0 + 1];
			J.imp[(i0 + 1) * K + k] = dot(*w, su) * start[i0] + dot(*w, sw) * start[i0 + 1];
		}
		jst3(J.u, k, K, *u);
		jst3(J.w, k, K, *w);
	};
	if (joint_is_inequality(kind)) {
		ja[0][0] = il_a.x; ja[0][1] = il_a.y; ja[0][2] = il_a.z;
		ja[0][3] = lim_a.x; ja[0][4] = lim_a.y; ja[0][5] = lim_a.z;
		jb[0][0] = il_b.x; jb[0][1] = il_b.y; jb[0][2] = il_b.z;
		jb[0][3] = lim_b.x; jb[0][4] = lim_b.y; jb[0][5] = lim_b.z;
		pt[0] = ipt;
	} else if (kind == 1) {  // DistanceConstraint (ConstraintSolver.cpp:348-379), not in moving mode
		V3 rA = pa - ca, rB = pb - cb;
		V3 diff = pa - pb;
		V3 dd = normalize(diff);
		double d[3] = {dd.x, dd.y, dd.z}, nd[3] = {-dd.x, -dd.y, -dd.z}, t[3];
		put3(ja, 0, 0, d);
		row_cross(nd, rA, t);
		put3(ja, 0, 3, t);
		put3(jb, 0, 0, nd);
		row_cross(d, rB, t);
		put3(jb, 0, 3, t);
		double target = J.ax1[k];
		double error = dot(diff, diff) < 0.00001 ? target : target - dot(normalize(diff), diff);
		pt[0] = pcc * error;
	} else if (kind == 3) {  // BallSocketConstraint (:382-408)
		V3 rA = pa - ca, rB = pb - cb;
		put_iden(ja, 0, 0, false);
		put_cross(ja, 0, 3, rA, true);
		put_iden(jb, 0, 0, true);
		put_cross(jb, 0, 3, rB, false);
		V3 err = pb - pa;
		pt[0] = pcc * err.x;
		pt[1] = pcc * err.y;
		pt[2] = pcc * err.z;
	} else if (kind == 5) {  // HingeConstraint (:468-553)
		V3 axa = q_rotate(qa, jld3(J.ax1, k, K)), axb = q_rotate(qb, jld3(J.ax2, k, K));
		V3 u, w;
		basis(axa, 3, &u, &w);
		V3 nn = axb;
		V3 rA = pa - ca, rB = pa - cb;  // both arms from the b1 point (ConstraintSolver.cpp:469)
		put_iden(ja, 0, 0, false);
		put_cross(ja, 0, 3, rA, true);
		put_iden(jb, 0, 0, true);
		put_cross(jb, 0, 3, rB, false);
		double ud[3] = {u.x, u.y, u.z}, wd[3] = {w.x, w.y, w.z}, nd[3] = {nn.x, nn.y, nn.z}, t[3];
		row_cross(nd, u, t); put3(ja, 3, 3, t);
		row_cross(ud, nn, t); put3(jb, 3, 3, t);
		row_cross(nd, w, t); put3(ja, 4, 3, t);
		row_cross(wd, nn, t); put3(jb, 4, 3, t);
		V3 pc = (pb - pa) * pcc;
		pt[0] = pc.x;
		pt[1] = pc.y;
		pt[2] = pc.z;
		pt[3] = dot(u, nn) * rcc;
		pt[4] = dot(w, nn) * rcc;
	} else if (kind == 7) {  // SliderConstraint (:611-697); Slider passes the position coefficient for both (PersistentConstraint.cpp:297)
		V3 u, w;
		basis(normalize(q_rotate(qa, jld3(J.ax1, k, K))), 0, &u, &w);
		V3 rA = pa - ca, rB = pb - cb;
		V3 pe = pa - pb;
		double ud[3] = {u.x, u.y, u.z}, wd[3] = {w.x, w.y, w.z}, nu[3] = {-u.x, -u.y, -u.z}, nw[3] = {-w.x, -w.y, -w.z};
		double ped[3] = {pe.x, pe.y, pe.z}, t1[3], t2[3];
		put3(ja, 0, 0, ud);
		row_cross(nu, rA, t1); row_cross(ped, u, t2);
		for (int c = 0; c < 3; c++) ja[0][3 + c] = t1[c] - t2[c];
		put3(jb, 0, 0, nu);
		row_cross(ud, rB, t1); put3(jb, 0, 3, t1);
		put3(ja, 1, 0, wd);
		row_cross(nw, rA, t1); row_cross(ped, w, t2);
		for (int c = 0; c < 3; c++) ja[1][3 + c] = t1[c] - t2[c];
		put3(jb, 1, 0, nw);
		row_cross(wd, rB, t1); put3(jb, 1, 3, t1);
		put_iden(ja, 2, 3, false);
		put_iden(jb, 2, 3, true);
		V3 rc = rot_correct(qa, qb, pcc);
		pt[0] = -dot(pe, u) * pcc;
		pt[1] = -dot(pe, w) * pcc;
		pt[2] = rc.x;
		pt[3] = rc.y;
		pt[4] = rc.z;
	} else if (kind == 4) {  // SlidingHingeConstraint (:781-857)
		V3 axa = q_rotate(qa, jld3(J.ax1, k, K)), axb = q_rotate(qb, jld3(J.ax2, k, K));
		V3 u, w;
		basis(axb, 2, &u, &w);
		V3 nn = axa;
		V3 rA = pa - ca, rB = pb - cb;
		V3 pd = pa - pb;
		double ud[3] = {u.x, u.y, u.z}, wd[3] = {w.x, w.y, w.z}, nu[3] = {-u.x, -u.y, -u.z}, nw[3] = {-w.x, -w.y, -w.z};
		double nnn[3] = {-nn.x, -nn.y, -nn.z}, pdd[3] = {pd.x, pd.y, pd.z}, t1[3], t2[3];
		put3(ja, 0, 0, ud);
		row_cross(nu, rA, t1); put3(ja, 0, 3, t1);
		put3(jb, 0, 0, nu);
		row_cross(ud, rB, t1); row_cross(pdd, u, t2);
		for (int c = 0; c < 3; c++) jb[0][3 + c] = t1[c] - t2[c];
		put3(ja, 1, 0, wd);
		row_cross(nw, rA, t1); put3(ja, 1, 3, t1);
		put3(jb, 1, 0, nw);
		row_cross(wd, rB, t1); row_cross(pdd, w, t2);
		for (int c = 0; c < 3; c++) jb[1][3 + c] = t1[c] - t2[c];
		row_cross(nu, nn, t1); put3(ja, 2, 3, t1);
		row_cross(nnn, u, t1); put3(jb, 2, 3, t1);
		row_cross(nw, nn, t1); put3(ja, 3, 3, t1);
		row_cross(nnn, w, t1); put3(jb, 3, 3, t1);
		pt[0] = -dot(u, pd) * pcc;
		pt[1] = -dot(w, pd) * pcc;
		pt[2] = -dot(u, nn) * rcc;
		pt[3] = -dot(w, nn) * rcc;
	} else {  // 6: WeldConstraint (:861-897)
		V3 rA = pa - ca, rB = pb - cb;
		put_iden(ja, 0, 0, false);
		put_cross(ja, 0, 3, rA, true);
		put_iden(ja, 3, 3, false);
		put_iden(jb, 0, 0, true);
		put_cross(jb, 0, 3, rB, false);
		put_iden(jb, 3, 3, true);
		V3 pc = (pb - pa) * pcc;
		V3 rc = rot_correct(qa, qb, rcc);
		pt[0] = pc.x; pt[1] = pc.y; pt[2] = pc.z;
		pt[3] = rc.x; pt[4] = rc.y; pt[5] = rc.z;
	}
	double ita[6][CP_JMAX], itb[6][CP_JMAX];
	for (int r = 0; r < 6; r++)
		for (int c = 0; c < n; c++) {
			double s1 = 0, s2 = 0;
			for (int q = 0; q < 6; q++) { s1 += Ma[r][q] * ja[c][q]; s2 += Mb[r][q] * jb[c][q]; }
			ita[r][c] = s1;
			itb[r][c] = s2;
		}
	double itv[CP_JMAX * CP_JMAX], inv[CP_JMAX * CP_JMAX];
	for (int r = 0; r < n; r++)
		for (int c = 0; c < n; c++) {
			double s1 = 0, s2 = 0;
			for (int q = 0; q < 6; q++) { s1 += ja[r][q] * ita[q][c]; s2 += jb[r][q] * itb[q][c]; }
			itv[r * n + c] = s1 + s2;
		}
	for (int r = 0; r < n; r++) itv[r * n + r] *= (1 + cfm);  // applyCFM (ConstraintSolver.cpp:919-926)
	inverse_n(n, itv, inv);
	for (int r = 0; r < n; r++)
		for (int c = 0; c < 6; c++) {
			R[CP_J_JA + r * 6 + c] = ja[r][c];
			R[CP_J_JB + r * 6 + c] = jb[r][c];
		}
	for (int r = 0; r < 6; r++)
		for (int c = 0; c < n; c++) {
			R[CP_J_ITA + r * CP_JMAX + c] = ita[r][c];
			R[CP_J_ITB + r * CP_JMAX + c] = itb[r][c];
		}
	for (int r = 0; r < n * n; r++) R[CP_J_INV + r] = inv[r];
	for (int c = 0; c < n; c++) {
		J.pimp[c * K + k] = 0;
		J.himp[c * K + k] = 0;
	}
}

// Joints k0..k1-1 as their constructors leave them: a motor's angle measured
// from the bodies (Motor ctor, PersistentConstraint.cpp:17-22) and a piston's
// position from its two attach points.  The Piston constructor (:139-141)
// looks b2's point key up in b1's tracked points; that value is replaced at
// the first update and is seen only by a position query before the first step.
__global__ __launch_bounds__(64) void k_joint_init(Bodies b, Joints J, int k0, int k1) {
	const int k = k0 + blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= k1) return;
	const int N = b.n, K = J.n, kind = J.kind[k];
	if (kind < 10) return;
	const int ia = J.a[k], ib = J.b[k];
	Q qa{b.q[ia], b.q[N + ia], b.q[2 * N + ia], b.q[3 * N + ia]};
	Q qb{b.q[ib], b.q[N + ib], b.q[2 * N + ib], b.q[3 * N + ib]};
	if (kind == 13) {  // addSpring (PhysicsEngine.cpp:978): a negative length takes the current distance
		const double given = J.ax1[2 * K + k];
		V3 pa = jld3(b.com, ia, N) + q_rotate(qa, jld3(J.tp1, k, K));
		V3 pb = jld3(b.com, ib, N) + q_rotate(qb, jld3(J.tp2, k, K));
		J.u[k] = given >= 0 ? given : mag(pa - pb);
		return;
	}
	if (kind == 10) {
		V3 z{0, 0, 0};
		J.u[k] = motor_position(J, k, qa, qb, q_rotate(qa, jld3(J.ax1, k, K)), z, z, 0, 0);
		return;
	}
	V3 pa = jld3(b.com, ia, N) + q_rotate(qa, jld3(J.tp1, k, K));
	V3 pb = jld3(b.com, ib, N) + q_rotate(qb, jld3(J.tp2, k, K));
	V3 axis = q_rotate(qa, jld3(J.ax1, k, K));
	J.u[(kind == 11 ? 0 : K) + k] = dot(pb, axis) - dot(pa, axis);
}

// mode 0: every joint except holonomic ones at the first substep; 1: holonomic
// joints with re-projection (end of step); 2: holonomic joints as stored
// (state loaded, or systems just created).
// A joint's rows take its own CFM when it has one (cfm.getCFMValue(global_cfm),
// PersistentConstraint.cpp:217-329); springs have no rows.
__global__ __launch_bounds__(64) void k_joint_rows(Bodies b, Joints J, Params P, double sub_dt, int mode, int first) {
	int k = blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= J.n || J.kind[k] == 13) return;
	bool in_sys = J.sys[k] >= 0 && P.hol_iters > 0;
	const double own = J.cfm[k];
	const double cfm = isnan(own) ? P.cfm : own;
	if (mode == 0) {
		if (first && in_sys) return;
		build_joint(b, J, k, sub_dt, cfm, true);
	} else {
		if (!in_sys) return;
		build_joint(b, J, k, sub_dt, cfm, mode == 1);
	}
}

// ------------------------------------------------------------------- springs
// RigidBody::applyImpulse (RigidBody.cpp:60-66) on a body whose inverse mass
// and world inverse inertia are im, it; only dynamic bodies are written (the
// reference adds zero products to the others).
__device__ __forceinline__ void spring_impulse(const Bodies& b, int i, bool dyn, double im, const M3& it, V3 com, V3 imp, V3 pos) {
	V3 torque = cross(pos - com, imp);
	V3 dvel = im * imp, dang = it * torque;
	if (!dyn) return;
	const int N = b.n;
	b.vel[i] += dvel.x; b.vel[N + i] += dvel.y; b.vel[2 * N + i] += dvel.z;
	b.angv[i] += dang.x; b.angv[N + i] += dang.y; b.angv[2 * N + i] += dang.z;
}

// Spring forces at a substep's start (PhysicsEngine.cpp:1313-1355): one
// thread per group of springs that share dynamic bodies, its springs in the
// constraint graph's edge order (nodes in node-map order, a node's edges by
// the other body's id, an edge's springs in creation order).  Each spring
// pushes both bodies along the attach-point direction by stiffness times the
// stretch, then damps their relative velocity along it, at most to a
// standstill.  getVelOfPoint (RigidBody.cpp:153-155) is taken at the other
// body's attach point, as the reference takes it.
__global__ __launch_bounds__(64) void k_springs(Bodies b, Joints J, int ngroups, const int* goff, const int* list, double dt) {
	const int g = blockIdx.x * blockDim.x + threadIdx.x;
	if (g >= ngroups) return;
	const int N = b.n, K = J.n;
	for (int x = goff[g]; x < goff[g + 1]; x++) {
		const int k = list[x];
		const int i1 = J.a[k], i2 = J.b[k];
		const bool d1 = b.movement[i1] == 0, d2 = b.movement[i2] == 0;
		const double im1 = d1 ? 1.0 / b.mass[i1] : 0, im2 = d2 ? 1.0 / b.mass[i2] : 0;  // getInvMass (RigidBody.cpp:161-163)
		M3 it1, it2;  // getInvTensor (:169-171)
		for (int r = 0; r < 3; r++)
			for (int c = 0; c < 3; c++) {
				it1.m[r][c] = d1 ? b.Iinv[(3 * r + c) * N + i1] : 0.0;
				it2.m[r][c] = d2 ? b.Iinv[(3 * r + c) * N + i2] : 0.0;
			}
		const V3 c1 = jld3(b.com, i1, N), c2 = jld3(b.com, i2, N);
		const Q q1{b.q[i1], b.q[N + i1], b.q[2 * N + i1], b.q[3 * N + i1]};
		const Q q2{b.q[i2], b.q[N + i2], b.q[2 * N + i2], b.q[3 * N + i2]};
		const double stiffness = J.ax1[k], damping = J.ax1[K + k], rest = J.u[k];
		const V3 p1 = c1 + q_rotate(q1, jld3(J.tp1, k, K)), p2 = c2 + q_rotate(q2, jld3(J.tp2, k, K));
		const double distance = mag(p2 - p1);
		const V3 dir = (distance != 0) ? normalize(p2 - p1) : V3{0, -1, 0};
		const double expand = stiffness * (distance - rest);
		spring_impulse(b, i1, d1, im1, it1, c1, dir * expand * dt, p1);
		spring_impulse(b, i2, d2, im2, it2, c2, -dir * expand * dt, p2);
		const V3 rA = p1 - c1, rB = p2 - c2;
		const V3 rdA = it1 * cross(dir, rA), rdB = it2 * cross(dir, rB);
		const double sens = 1.0 / (im1 + im2 + dot(cross(rA, rdA), dir) + dot(cross(rB, rdB), dir));
		const V3 v1 = jld3(b.vel, i1, N), w1 = jld3(b.angv, i1, N), v2 = jld3(b.vel, i2, N), w2 = jld3(b.angv, i2, N);
		const double velocity = dot(v2 + cross(w2, p1 - c2), dir) - dot(v1 + cross(w1, p2 - c1), dir);
		double dampen = damping * velocity;
		const double max_dampen = velocity * sens / dt;
		if (velocity > 0 && dampen > max_dampen) dampen = max_dampen;
		else if (velocity < 0 && dampen < max_dampen) dampen = max_dampen;
		spring_impulse(b, i1, d1, im1, it1, c1, dir * dampen * dt, p1);
		spring_impulse(b, i2, d2, im2, it2, c2, -dir * dampen * dt, p2);
	}
}

// ------------------------------------------------------------- holonomic factor
__device__ __forceinline__ int hs_first_below(const HSysDev& H, int s, int col) {  // :616-622
	for (int row = col + 1; row < H.nc[s]; row++) {
		int l = hs_loc(H, s, row, col);
		if (l != -1) return l;
	}
	return -1;
}
// computeBlockInitialValue (:799-825, :714-744), row joint c1, column joint c2.
__device__ void hs_block_init(const Joints& J, double cfm, int c1, int c2, double* t) {
	int n1 = J.type[c1], n2 = J.type[c2];
	const double* R1 = J.row + (size_t)c1 * CP_JRS;
	const double* R2 = J.row + (size_t)c2 * CP_JRS;
	if (c1 == c2) {
		for (int r = 0; r < n1; r++)
			for (int c = 0; c < n1; c++) {
				double s1 = 0, s2 = 0;
				for (int q = 0; q < 6; q++) {
					s1 += R1[CP_J_JA + r * 6 + q] * R1[CP_J_ITA + q * CP_JMAX + c];
					s2 += R1[CP_J_JB + r * 6 + q] * R1[CP_J_ITB + q * CP_JMAX + c];
				}
				t[r * n1 + c] = s1 + s2;
			}
		for (int i = 0; i < n1; i++) t[i + n1 * i] *= (1 + cfm);
		return;
	}
	int jac = -1, it = -1;
	if (J.a[c1] == J.a[c2]) { jac = CP_J_JA; it = CP_J_ITA; }
	else if (J.a[c1] == J.b[c2]) { jac = CP_J_JA; it = CP_J_ITB; }
	else if (J.b[c1] == J.a[c2]) { jac = CP_J_JB; it = CP_J_ITA; }
	else if (J.b[c1] == J.b[c2]) { jac = CP_J_JB; it = CP_J_ITB; }
	if (jac < 0) {
		for (int i = 0; i < n1 * n2; i++) t[i] = 0;
		return;
	}
	for (int i = 0; i < n1 * n2; i++) {  // matMult (:701-712)
		int row = i / n2, col = i % n2;
		t[i] = 0;
		for (int q = 0; q < 6; q++) t[i] += R1[jac + row * 6 + q] * R2[it + q * CP_JMAX + col];
	}
}

// HolonomicSystem::computeInverse (:459-543), one thread per system.
__global__ __launch_bounds__(64) void k_hsys_factor(Joints J, HSysDev H, double cfm) {
	int s = blockIdx.x * blockDim.x + threadIdx.x;
	if (s >= H.nsys) return;
	const int nc = H.nc[s];
	const int* ord = H.order + H.ord_off[s];
	double* buf = H.buf + H.buf_off[s];
	double* dbuf = H.dbuf + H.dbuf_off[s];
	for (int col = 0; col < nc; col++)
		for (int row = col; row < nc; row++) {
			int l = hs_loc(H, s, row, col);
			if (l == -1) continue;
			hs_block_init(J, cfm, ord[row], ord[col], buf + l);
		}
	for (int col = 0; col < nc; col++) {
		double* diag = buf + hs_loc(H, s, col, col);
		int bw = J.type[ord[col]];
		double tmp[36];
		inverse_n(bw, diag, tmp);
		for (int q = 0; q < bw * bw; q++) diag[q] = tmp[q];
		if (col + 1 == nc) continue;
		int first = hs_first_below(H, s, col);
		for (int row = col + 1; row < nc; row++) {
			int l = hs_loc(H, s, row, col);
			if (l == -1) continue;
			double* tgt = dbuf + (l - first);
			int bh = J.type[ord[row]];
			const double* src = buf + l;
			for (int r = 0; r < bh; r++)
				for (int c = 0; c < bw; c++) tgt[c + bw * r] = dotp(bw, src + bw * r, diag + bw * c);
		}
		for (int ldc = col + 1; ldc < nc; ldc++)
			for (int ldr = ldc; ldr < nc; ldr++) {
				int l = hs_loc(H, s, ldr, col);
				int lc = hs_loc(H, s, ldc, col);
				if (l == -1 || lc == -1) continue;
				const double* src = buf + l;
				const double* lsrc = dbuf + (lc - first);
				double* tgt = buf + hs_loc(H, s, ldr, ldc);
				int w2 = J.type[ord[ldc]], h2 = J.type[ord[ldr]];
				for (int r = 0; r < h2; r++)
					for (int c = 0; c < w2; c++) tgt[r * w2 + c] -= dotp(bw, src + r * bw, lsrc + c * bw);
			}
		if (first == -1) continue;
		int total = hs_loc(H, s, col + 1, col + 1) - first;
		for (int q = 0; q < total; q++) buf[first + q] = dbuf[q];
	}
}

// One element (r, c) of block (c1 row joint, c2 column joint) of the initial
// values: computeBlockInitialValue's arithmetic for that element alone.
__device__ double hs_block_elem(const Joints& J, double cfm, int c1, int c2, int r, int c) {
	const double* R1 = J.row + (size_t)c1 * CP_JRS;
	const double* R2 = J.row + (size_t)c2 * CP_JRS;
	if (c1 == c2) {
		double s1 = 0, s2 = 0;
		for (int q = 0; q < 6; q++) {
			s1 += R1[CP_J_JA + r * 6 + q] * R1[CP_J_ITA + q * CP_JMAX + c];
			s2 += R1[CP_J_JB + r * 6 + q] * R1[CP_J_ITB + q * CP_JMAX + c];
		}
		double t = s1 + s2;
		if (r == c) t *= (1 + cfm);
		return t;
	}
	int jac = -1, it = -1;
	if (J.a[c1] == J.a[c2]) { jac = CP_J_JA; it = CP_J_ITA; }
	else if (J.a[c1] == J.b[c2]) { jac = CP_J_JA; it = CP_J_ITB; }
	else if (J.b[c1] == J.a[c2]) { jac = CP_J_JB; it = CP_J_ITA; }
	else if (J.b[c1] == J.b[c2]) { jac = CP_J_JB; it = CP_J_ITB; }
	if (jac < 0) return 0;
	double t = 0;
	for (int q = 0; q < 6; q++) t += R1[jac + r * 6 + q] * R2[it + q * CP_JMAX + c];
	return t;
}

// HolonomicSystem::computeInverse (:459-543), one wave per system on its block
// buffer and column scratch staged in LDS: every step's block elements (the
// initial values, the lower-triangle blocks of a column, the updates of the
// blocks below and right of it, the copy back) run side by side, each element
// with the reference's operation order, and each diagonal block is inverted
// by one lane, so the factor equals k_hsys_factor's bit for bit.  A column's
// blocks below the diagonal lie contiguously in row order from its first one,
// each row-major.
__global__ __launch_bounds__(64) void k_hsys_factor_wave(Joints J, HSysDev H, double cfm) {
	extern __shared__ double sm[];
	const int s = blockIdx.x, lane = threadIdx.x;
	const int nc = H.nc[s];
	const int* ord = H.order + H.ord_off[s];
	const int cap = H.buf_off[s + 1] - H.buf_off[s];
	double* buf = sm;
	double* dbuf = sm + cap;
	const int* cs_ptr = H.cs_ptr + H.cp_off[s];  // per column: its blocks below the diagonal, rows ascending
	auto sync = [] {
		__builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
		__builtin_amdgcn_wave_barrier();
	};
	// initial values of every non-empty block, column by column
	for (int col = 0; col < nc; col++) {
		const int bw = J.type[ord[col]];
		const int dl = H.dg_loc[H.ord_off[s] + col];
		for (int e = lane; e < bw * bw; e += 64) buf[dl + e] = hs_block_elem(J, cfm, ord[col], ord[col], e / bw, e % bw);
		for (int x = cs_ptr[col]; x < cs_ptr[col + 1]; x++) {
			const int row = H.cs_row[x], l = H.cs_loc[x], bh = J.type[ord[row]];
			for (int e = lane; e < bh * bw; e += 64) buf[l + e] = hs_block_elem(J, cfm, ord[row], ord[col], e / bw, e % bw);
		}
	}
	sync();
	for (int col = 0; col < nc; col++) {
		const int bw = J.type[ord[col]];
		double* diag = buf + H.dg_loc[H.ord_off[s] + col];
		if (lane == 0) {
			double tmp[36];
			inverse_n(bw, diag, tmp);
			for (int q = 0; q < bw * bw; q++) diag[q] = tmp[q];
		}
		sync();
		const int x0 = cs_ptr[col], x1 = cs_ptr[col + 1];
		if (x0 == x1) continue;
		const int first = H.cs_loc[x0];
		const int total = hs_loc(H, s, col + 1, col + 1) - first;
		// lower-triangle blocks of the column into the scratch
		for (int e = lane; e < total; e += 64) {
			int x = x0;
			while (x + 1 < x1 && H.cs_loc[x + 1] - first <= e) x++;
			const int rel = e - (H.cs_loc[x] - first), r = rel / bw, c = rel - r * bw;
			dbuf[e] = dotp(bw, buf + H.cs_loc[x] + bw * r, diag + bw * c);
		}
		sync();
		// the blocks below and right of the column
		for (int xc = x0; xc < x1; xc++)
			for (int xr = xc; xr < x1; xr++) {
				const int ldc = H.cs_row[xc], ldr = H.cs_row[xr];
				const int w2 = J.type[ord[ldc]], h2 = J.type[ord[ldr]];
				const double* src = buf + H.cs_loc[xr];
				const double* lsrc = dbuf + (H.cs_loc[xc] - first);
				double* tgt = buf + hs_loc(H, s, ldr, ldc);
				for (int e = lane; e < h2 * w2; e += 64) {
					const int r = e / w2, c = e - r * w2;
					tgt[r * w2 + c] -= dotp(bw, src + r * bw, lsrc + c * bw);
				}
			}
		sync();
		for (int e = lane; e < total; e += 64) buf[first + e] = dbuf[e];
		sync();
	}
	double* out = H.buf + H.buf_off[s];
	for (int e = lane; e < cap; e += 64) out[e] = buf[e];
}

static inline int nblk(int n, int t) { return (n + t - 1) / t; }

void launch_joint_init(const Bodies& b, const Joints& J, int k0, int k1, hipStream_t s) {
	if (k1 > k0) hipLaunchKernelGGL(k_joint_init, dim3(nblk(k1 - k0, 64)), dim3(64), 0, s, b, J, k0, k1);
}
void launch_springs(const Bodies& b, const Joints& J, int ngroups, const int* goff, const int* list, double sub_dt, hipStream_t s) {
	if (ngroups) hipLaunchKernelGGL(k_springs, dim3(nblk(ngroups, 64)), dim3(64), 0, s, b, J, ngroups, goff, list, sub_dt);
}
void launch_joint_rows(const Bodies& b, const Joints& J, const Params& P, double sub_dt, int mode, int first, hipStream_t s) {
	if (J.n) hipLaunchKernelGGL(k_joint_rows, dim3(nblk(J.n, 64)), dim3(64), 0, s, b, J, P, sub_dt, mode, first);
}
void launch_hsys_factor(const Joints& J, const HSysDev& H, double cfm, hipStream_t s) {
	if (!H.nsys) return;
	const size_t lds = (size_t)H.lds_doubles * sizeof(double);
	if (lds > 0 && lds <= 64 * 1024) {
		static bool attr = false;
		if (!attr) {
			(void)hipFuncSetAttribute((const void*)k_hsys_factor_wave, hipFuncAttributeMaxDynamicSharedMemorySize, 64 * 1024);
			attr = true;
		}
		hipLaunchKernelGGL(k_hsys_factor_wave, dim3(H.nsys), dim3(64), lds, s, J, H, cfm);
	} else
		hipLaunchKernelGGL(k_hsys_factor, dim3(nblk(H.nsys, 64)), dim3(64), 0, s, J, H, cfm);
}

}  // namespace cpx
