// Per-body integration shared by the integration kernel and the fused
// substep kernels: SoA accessors, RigidBody::rotateWhileApplyingGyroAccel and
// the substep position / orientation / inertia update, FP64 in the
// reference's operation order (compile with -ffp-contract=off).
#pragma once
#include "cp_internal.h"

namespace cpx {

#define LD3(a, i, n) V3{(a)[(i)], (a)[(n) + (i)], (a)[2 * (n) + (i)]}
#define ST3(a, i, n, v) do { (a)[(i)] = (v).x; (a)[(n) + (i)] = (v).y; (a)[2 * (n) + (i)] = (v).z; } while (0)

__device__ __forceinline__ Q ldq(const double* a, int i, int n) { return Q{a[i], a[n + i], a[2 * n + i], a[3 * n + i]}; }
__device__ __forceinline__ void stq(double* a, int i, int n, Q q) {
	a[i] = q.r; a[n + i] = q.i; a[2 * n + i] = q.j; a[3 * n + i] = q.k;
}
__device__ __forceinline__ M3 ldm(const double* a, int i, int n) {
	M3 m;
	for (int r = 0; r < 3; r++) for (int c = 0; c < 3; c++) m.m[r][c] = a[(3 * r + c) * n + i];
	return m;
}
__device__ __forceinline__ void stm(double* a, int i, int n, const M3& m) {
	for (int r = 0; r < 3; r++) for (int c = 0; c < 3; c++) a[(3 * r + c) * n + i] = m.m[r][c];
}

// RigidBody::rotateWhileApplyingGyroAccel (RigidBody.cpp:207-255), implicit
// Euler on the gyroscopic term with two Newton iterations per sub-tick.
static __device__ void gyro_rotate(V3& angv, Q& q, const M3& I_world, const M3& I_ref, double dt, int n_itr) {
	if (mag2(angv) == 0) return;
	M3 T = I_world;
	Q o = q;
	V3 av = angv;
	for (int it = 0; it < n_itr; it++) {
		V3 w = av;
		double ts = dt / n_itr;
		for (int s = 0; s < 2; s++) {
			V3 f = ts * cross(w, T * w) - T * (av - w);
			M3 jac = T + ts * (cross_mat(w) * T - cross_mat(T * w));
			if (fabs(det(jac)) > 0.00000000001) w -= inverse(jac) * f;
		}
		av = w;
		o = q_mul(q_axis_angle(ts * mag(av), av), o);
		if (it + 1 < n_itr) {
			M3 r = q_rot(o), rc = q_rot(q_conj(o));
			T = r * I_ref * rc;
		}
	}
	q = o;
	angv = av;
}

// update_positions (PhysicsEngine.cpp:390-410) followed by
// RigidBody::updateInertiaTensor (RigidBody.cpp:270-279) for dynamic body i.
// iinv32 (optional): also stores the new world inverse inertia's upper
// triangle in FP32 there ([n][8]: xx xy xz yy yz zz, 2 pad), the values the
// FP32 solve reads.
static __device__ void integrate_body(const Bodies& b, const Params& p, int i, float* iinv32 = nullptr) {
	const int n = b.n;
	V3 com = LD3(b.com, i, n), vel = LD3(b.vel, i, n), angv = LD3(b.angv, i, n);
	V3 pvel = LD3(b.pvel, i, n), pang = LD3(b.pang, i, n);
	Q q = ldq(b.q, i, n);
	ST3(b.prev_com, i, n, com);
	stq(b.prev_q, i, n, q);
	com += (vel + pvel) * p.sub_dt;
	if (mag2(angv) != 0) {
		M3 Iw = ldm(b.I, i, n), Ir = ldm(b.Iref, i, n);
		gyro_rotate(angv, q, Iw, Ir, p.sub_dt, p.gyro_ticks);
		q = q_mul(q_axis_angle(p.sub_dt * mag(pang), pang), q);
	}
	V3 z{0, 0, 0};
	ST3(b.pvel, i, n, z);
	ST3(b.pang, i, n, z);
	q = q_normalize(q);
	M3 r = q_rot(q), rc = q_rot(q_conj(q));
	M3 Iw = r * ldm(b.Iref, i, n) * rc;
	M3 Iwi = r * ldm(b.Irefinv, i, n) * rc;
	ST3(b.com, i, n, com);
	ST3(b.angv, i, n, angv);
	stq(b.q, i, n, q);
	stm(b.I, i, n, Iw);
	stm(b.Iinv, i, n, Iwi);
	if (iinv32) {
		float4* o = reinterpret_cast<float4*>(iinv32 + 8 * (size_t)i);
		o[0] = float4{(float)Iwi.m[0][0], (float)Iwi.m[0][1], (float)Iwi.m[0][2], (float)Iwi.m[1][1]};
		o[1] = float4{(float)Iwi.m[1][2], (float)Iwi.m[2][2], 0.0f, 0.0f};
	}
}

}  // namespace cpx
