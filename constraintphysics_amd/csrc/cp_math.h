// Small FP64 vector / quaternion / 3x3 algebra shared by the host runtime and
// the gfx950 kernels.  Operation order follows ConstraintPhysics' mthz library
// (Math/src/Vec3.h, Quaternion.h, Mat3.h, NMat.h) so that, compiled without
// FMA contraction, every result is the reference's IEEE result.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#define CPX_HD __host__ __device__ __forceinline__

namespace cpx {

struct V3 {
	double x, y, z;
};
CPX_HD V3 v3(double a, double b, double c) { return V3{a, b, c}; }
CPX_HD V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
CPX_HD V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
CPX_HD V3 operator-(V3 a) { return V3{-a.x, -a.y, -a.z}; }
CPX_HD V3 operator*(V3 a, double d) { return V3{a.x * d, a.y * d, a.z * d}; }  // Vec3.cpp:4-9 (both argument orders)
CPX_HD V3 operator*(double d, V3 a) { return V3{a.x * d, a.y * d, a.z * d}; }
CPX_HD void operator+=(V3& a, V3 b) { a.x += b.x; a.y += b.y; a.z += b.z; }
CPX_HD void operator-=(V3& a, V3 b) { a.x -= b.x; a.y -= b.y; a.z -= b.z; }
CPX_HD double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
CPX_HD V3 cross(V3 a, V3 b) { return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
CPX_HD double mag2(V3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
CPX_HD double mag(V3 a) { return sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
CPX_HD V3 normalize(V3 a) {  // Vec3.h:38-41
	double m = mag(a);
	return (m == 0) ? V3{0, 0, 0} : V3{a.x / m, a.y / m, a.z / m};
}
CPX_HD bool neq(V3 a, V3 b) { return a.x != b.x || a.y != b.y || a.z != b.z; }
// Vec3::getPerpendicularBasis (Vec3.h:53-57).
CPX_HD void perp_basis(V3 n, V3* u, V3* w) {
	double d1 = n.x * 1.0 + n.y * 0.0 + n.z * 0.0;
	double d2 = n.x * 0.0 + n.y * 1.0 + n.z * 0.0;
	*u = (fabs(d1) < fabs(d2)) ? normalize(cross(n, V3{1, 0, 0})) : normalize(cross(n, V3{0, 1, 0}));
	*w = normalize(cross(n, *u));
}

struct M3 {
	double m[3][3];
};
CPX_HD M3 m3_zero() {
	M3 o;
	for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) o.m[i][j] = 0;
	return o;
}
CPX_HD V3 operator*(const M3& a, V3 w) {  // Mat3.h:70-74
	return V3{a.m[0][0] * w.x + a.m[0][1] * w.y + a.m[0][2] * w.z, a.m[1][0] * w.x + a.m[1][1] * w.y + a.m[1][2] * w.z,
	          a.m[2][0] * w.x + a.m[2][1] * w.y + a.m[2][2] * w.z};
}
CPX_HD M3 operator*(const M3& a, const M3& b) {  // Mat3.h:95-106
	M3 o = m3_zero();
	for (int i = 0; i < 3; i++)
		for (int j = 0; j < 3; j++)
			for (int k = 0; k < 3; k++) o.m[i][j] += a.m[i][k] * b.m[k][j];
	return o;
}
CPX_HD M3 operator+(const M3& a, const M3& b) {
	M3 o;
	for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) o.m[i][j] = a.m[i][j] + b.m[i][j];
	return o;
}
CPX_HD M3 operator-(const M3& a, const M3& b) {
	M3 o;
	for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) o.m[i][j] = a.m[i][j] - b.m[i][j];
	return o;
}
CPX_HD M3 operator*(double d, const M3& a) {  // Mat3.cpp:4-12
	M3 o;
	for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) o.m[i][j] = d * a.m[i][j];
	return o;
}
CPX_HD double det(const M3& q) {  // Mat3.h:52-55
	double a = q.m[0][0], b = q.m[0][1], c = q.m[0][2], d = q.m[1][0], e = q.m[1][1], f = q.m[1][2], g = q.m[2][0], h = q.m[2][1],
	       i = q.m[2][2];
	return a * (e * i - f * h) - b * (d * i - f * g) + c * (d * h - e * g);
}
CPX_HD M3 inverse(const M3& q) {  // Mat3.h:57-68
	M3 o;
	double a = q.m[0][0], b = q.m[0][1], c = q.m[0][2], d = q.m[1][0], e = q.m[1][1], f = q.m[1][2], g = q.m[2][0], h = q.m[2][1],
	       i = q.m[2][2];
	double dt = a * (e * i - f * h) - b * (d * i - f * g) + c * (d * h - e * g);
	o.m[0][0] = (e * i - f * h); o.m[0][1] = (c * h - b * i); o.m[0][2] = (b * f - c * e);
	o.m[1][0] = (f * g - d * i); o.m[1][1] = (a * i - c * g); o.m[1][2] = (c * d - a * f);
	o.m[2][0] = (d * h - e * g); o.m[2][1] = (b * g - a * h); o.m[2][2] = (a * e - b * d);
	for (int r = 0; r < 3; r++) for (int s = 0; s < 3; s++) o.m[r][s] /= dt;
	return o;
}
CPX_HD M3 cross_mat(V3 w) {  // Mat3.h:43-49
	M3 o;
	o.m[0][0] = 0; o.m[0][1] = -w.z; o.m[0][2] = w.y;
	o.m[1][0] = w.z; o.m[1][1] = 0; o.m[1][2] = -w.x;
	o.m[2][0] = -w.y; o.m[2][1] = w.x; o.m[2][2] = 0;
	return o;
}

struct Q {
	double r, i, j, k;
};
CPX_HD Q q_axis_angle(double theta, V3 axis) {  // Quaternion.h:26-33
	axis = normalize(axis);
	Q q;
	q.r = cos(theta / 2.0);
	double s = sin(theta / 2.0);
	q.i = axis.x * s; q.j = axis.y * s; q.k = axis.z * s;
	return q;
}
CPX_HD V3 q_rotate(Q q, V3 v) {  // Quaternion.h:51-55
	double r = q.r, i = q.i, j = q.j, k = q.k;
	return V3{v.x * (r * r + i * i - j * j - k * k) + v.y * (2 * i * j - 2 * r * k) + v.z * (2 * i * k + 2 * r * j),
	          v.x * (2 * i * j + 2 * r * k) + v.y * (r * r - i * i + j * j - k * k) + v.z * (2 * j * k - 2 * r * i),
	          v.x * (2 * i * k - 2 * r * j) + v.y * (2 * j * k + 2 * r * i) + v.z * (r * r - i * i - j * j + k * k)};
}
CPX_HD M3 q_rot(Q q) {  // Quaternion.h:61-67
	double r = q.r, i = q.i, j = q.j, k = q.k;
	M3 m;
	m.m[0][0] = r * r + i * i - j * j - k * k; m.m[0][1] = 2 * i * j - 2 * r * k; m.m[0][2] = 2 * i * k + 2 * r * j;
	m.m[1][0] = 2 * i * j + 2 * r * k; m.m[1][1] = r * r - i * i + j * j - k * k; m.m[1][2] = 2 * j * k - 2 * r * i;
	m.m[2][0] = 2 * i * k - 2 * r * j; m.m[2][1] = 2 * j * k + 2 * r * i; m.m[2][2] = r * r - i * i - j * j + k * k;
	return m;
}
CPX_HD Q q_conj(Q q) { return Q{q.r, -q.i, -q.j, -q.k}; }
CPX_HD Q q_mul(Q a, Q q) {  // Quaternion.h:73-78
	return Q{a.r * q.r - a.i * q.i - a.j * q.j - a.k * q.k, a.r * q.i + a.i * q.r + a.j * q.k - a.k * q.j,
	         a.r * q.j - a.i * q.k + a.j * q.r + a.k * q.i, a.r * q.k + a.i * q.j - a.j * q.i + a.k * q.r};
}
CPX_HD Q q_normalize(Q q) {  // Quaternion.h:88-91
	double m = sqrt(q.r * q.r + q.i * q.i + q.j * q.j + q.k * q.k);
	return Q{q.r / m, q.i / m, q.j / m, q.k / m};
}
CPX_HD bool q_neq(Q a, Q b) { return a.r != b.r || a.i != b.i || a.j != b.j || a.k != b.k; }

// NVec::isZero (NVec.h:12-19, CUTOFF 1e-8).
CPX_HD bool is_zero1(double a) { return !(fabs(a) > 0.00000001); }
CPX_HD bool is_zero2(double a, double b) { return !(fabs(a) > 0.00000001 || fabs(b) > 0.00000001); }

}  // namespace cpx
