// Broadphase: sort-and-sweep over world AABBs.
//
// The reference's dynamic AABB tree (AABB_Tree.h:153-224) reports every pair
// of leaves whose TRUE AABBs overlap (closed intervals, AABB.h:10-14); internal
// nodes are only conservative filters, so the reported set equals the
// brute-force overlap set.  Here: per scene, bodies are radix-sorted by
// AABB min.x (order-preserving u64 image of the double), the sorted bounds are
// gathered into SoA, and thread p sweeps forward from sorted position p while
// min.x[q] <= max.x[p], applying the reference's exact closed-interval test.
// Count pass, exclusive scan, write pass; the pair keys (lo << 32 | hi, global
// body indices) are then radix-sorted so the output order is deterministic.
//
// Roofline: HBM-bound.  Per body: read AABB 48 B (f64) + key/index 12 B, write
// sorted bounds 48 B; per emitted pair 8 B.
#include "cp_internal.h"

namespace cpx {

__global__ void k_bp_keys(Bodies b, uint64_t* keys, int* vals) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= b.nl) return;
	uint64_t u = (uint64_t)__double_as_longlong(b.aabb[i]);
	keys[i] = (u >> 63) ? ~u : (u | 0x8000000000000000ull);
	vals[i] = i;
}

// sb: sorted bounds of the b.nl visited bodies, stride b.nl.
__global__ void k_bp_gather(Bodies b, const int* sorted, double* sb) {
	int p = blockIdx.x * blockDim.x + threadIdx.x;
	const int m = b.nl;
	if (p >= m) return;
	int i = sorted[p];
	for (int c = 0; c < 6; c++) sb[c * m + p] = b.aabb[c * b.n + i];
}

// scene_end[p]: one past the last sorted position of p's scene.
__global__ void k_bp_sweep(int n, const int* sorted, const double* sb, const int* scene_of_body, const int* scene_body_off,
                           int* counts, const int* offs, uint64_t* out, int write) {
	int p = blockIdx.x * blockDim.x + threadIdx.x;
	if (p >= n) return;
	int i = sorted[p];
	int end = scene_body_off[scene_of_body[i] + 1];
	double amin0 = sb[p], amin1 = sb[n + p], amin2 = sb[2 * n + p];
	double amax0 = sb[3 * n + p], amax1 = sb[4 * n + p], amax2 = sb[5 * n + p];
	int cnt = 0;
	int base = write ? offs[p] : 0;
	for (int q = p + 1; q < end; q++) {
		double bmin0 = sb[q];
		if (bmin0 > amax0) break;
		double bmin1 = sb[n + q], bmin2 = sb[2 * n + q];
		double bmax0 = sb[3 * n + q], bmax1 = sb[4 * n + q], bmax2 = sb[5 * n + q];
		bool sep = amax0 < bmin0 || amin0 > bmax0 || amax1 < bmin1 || amin1 > bmax1 || amax2 < bmin2 || amin2 > bmax2;
		if (sep) continue;
		if (write) {
			int j = sorted[q];
			uint64_t lo = (uint64_t)(i < j ? i : j), hi = (uint64_t)(i < j ? j : i);
			out[base + cnt] = (lo << 32) | hi;
		}
		cnt++;
	}
	if (!write) counts[p] = cnt;
}

static inline int nblk(int n, int t) { return (n + t - 1) / t; }

void launch_bp_keys(const Bodies& b, uint64_t* keys, int* vals, hipStream_t s) {
	if (b.nl) hipLaunchKernelGGL(k_bp_keys, dim3(nblk(b.nl, 256)), dim3(256), 0, s, b, keys, vals);
}
void launch_bp_gather(const Bodies& b, const int* sorted, double* sb, hipStream_t s) {
	if (b.nl) hipLaunchKernelGGL(k_bp_gather, dim3(nblk(b.nl, 256)), dim3(256), 0, s, b, sorted, sb);
}
void launch_bp_sweep(int n, const int* sorted, const double* sb, const int* scene_of_body, const int* scene_body_off, int* counts,
                     const int* offs, uint64_t* out, int write, hipStream_t s) {
	if (n) hipLaunchKernelGGL(k_bp_sweep, dim3(nblk(n, 128)), dim3(128), 0, s, n, sorted, sb, scene_of_body, scene_body_off, counts, offs,
	                          out, write);
}

}  // namespace cpx
