// Device helpers shared by the joint-row, factor and PGS kernels: degreed
// constraint value / velocity change (ConstraintSolver.h:71-88) and the
// holonomic block solve (HolonomicBlockSolver.cpp:223-355), FP64 in the
// reference's operation order.
#pragma once
#include "cp_internal.h"

namespace cpx {

__device__ __forceinline__ int hs_loc(const HSysDev& H, int s, int row, int col) {  // getBlockBufferLocation (:603-614)
	int n = H.nc[s];
	return H.table[H.tab_off[s] + row - col + ((2 * n + 1) * col - col * col) / 2];
}
__device__ __forceinline__ double dotp(int n, const double* a, const double* b) {  // dotProd (:760-768)
	double o = 0;
	for (int i = 0; i < n; i++) o += a[i] * b[i];
	return o;
}

// J_a va + J_b vb of joint k (rows in R), va/vb strided by vs.
__device__ __forceinline__ void joint_value(const double* R, int n, const double* va, const double* vb, int vs, double* out) {
	for (int r = 0; r < n; r++) {
		double x = 0, y = 0;
		for (int q = 0; q < 6; q++) {
			x += va[q * vs] * R[CP_J_JA + r * 6 + q];
			y += vb[q * vs] * R[CP_J_JB + r * 6 + q];
		}
		out[r] = x + y;
	}
}
// va += M_a^-1 J_a^T imp, vb += M_b^-1 J_b^T imp (only dynamic bodies keep theirs).
__device__ __forceinline__ void joint_apply(const double* R, int n, const double* imp, double* va, double* vb, int vs, bool dyn_a, bool dyn_b) {
	double da[6], db[6];
	for (int r = 0; r < 6; r++) {
		da[r] = 0;
		db[r] = 0;
		for (int q = 0; q < n; q++) {
			da[r] += imp[q] * R[CP_J_ITA + r * CP_JMAX + q];
			db[r] += imp[q] * R[CP_J_ITB + r * CP_JMAX + q];
		}
	}
	for (int r = 0; r < 6; r++) {
		if (dyn_a) va[r * vs] += da[r];
		if (dyn_b) vb[r * vs] += db[r];
	}
}
__device__ __forceinline__ bool is_zero_n(const double* v, int n) {  // NVec::isZero (NVec.h:12-19)
	for (int i = 0; i < n; i++)
		if (fabs(v[i]) > 0.00000001) return false;
	return true;
}

// multInverseWithVector (:223-296) in place on v (degree entries); d is a
// scratch vector of the same length.  The reference zeroes a full-degree
// vector per column, fills the rows of the column's non-empty blocks and
// subtracts the whole vector; here only the non-empty blocks are visited (the
// skipped entries subtract an exact zero, which leaves every value as it is).
__device__ void hs_solve(const Joints& J, const HSysDev& H, int s, double* v, double* eff, double* d) {
	(void)eff;
	const int nc = H.nc[s];
	const int deg = H.degree[s];
	const int* ord = H.order + H.ord_off[s];
	const int* vl = H.vloc + H.ord_off[s];
	const int* cp = H.cs_ptr + H.cp_off[s];
	const int* dg = H.dg_loc + H.ord_off[s];
	const double* buf = H.buf + H.buf_off[s];
	for (int col = 0; col + 1 < nc; col++) {
		int bw = J.type[ord[col]];
		for (int e = cp[col]; e < cp[col + 1]; e++) {
			int row = H.cs_row[e], l = H.cs_loc[e];
			int bh = J.type[ord[row]];
			for (int i = 0; i < bh; i++) v[vl[row] + i] -= dotp(bw, buf + l + i * bw, v + vl[col]);
		}
	}
	for (int col = 0; col < nc; col++) {
		int bw = J.type[ord[col]];
		const double* m = buf + dg[col];
		for (int i = 0; i < bw; i++) d[vl[col] + i] = dotp(bw, m + i * bw, v + vl[col]);
	}
	for (int i = 0; i < deg; i++) v[i] = d[i];
	for (int col = nc - 2; col >= 0; col--) {
		double lt[6] = {0, 0, 0, 0, 0, 0};
		int bw = J.type[ord[col]];
		for (int e = cp[col]; e < cp[col + 1]; e++) {
			int row = H.cs_row[e], l = H.cs_loc[e];
			int bh = J.type[ord[row]];
			for (int i = 0; i < bw; i++) {
				double dp = 0;
				for (int q = 0; q < bh; q++) dp += v[vl[row] + q] * buf[l + i + bw * q];
				lt[i] += dp;
			}
		}
		for (int i = 0; i < bw; i++) v[vl[col] + i] -= lt[i];
	}
}

}  // namespace cpx
