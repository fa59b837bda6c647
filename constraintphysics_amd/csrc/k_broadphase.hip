// Broadphase: sort-and-sweep over world AABBs.
//
// The reference's dynamic AABB tree (AABB_Tree.h:153-224) reports every pair
// of leaves whose TRUE AABBs overlap (closed intervals, AABB.h:10-14); internal
// nodes are only conservative filters, so the reported set equals the
// brute-force overlap set.  Here: per scene, bodies are radix-sorted by
// AABB min.x (order-preserving u64 image of the double), the sorted bounds are
// gathered into SoA, and sorted position p is swept forward over q > p while
// min.x[q] <= max.x[p], applying the reference's exact closed-interval test.
// The sweep end of p is found by binary search and the range cut into chunks
// of BP_CH positions, one thread per chunk, so a body whose box spans the
// scene (a terrain mesh) costs no more than any other.  Count pass, exclusive
// scan, write pass; the pair keys (lo << 32 | hi, global body indices) are
// then radix-sorted so the output order is deterministic.
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

// sb: sorted bounds of the b.nl visited bodies, stride b.nl.  With
// skip_nonfinite, a box holding a NaN or an infinity is stored empty (min
// +inf, max -inf), so it overlaps nothing; the reference's test
// (AABB::intersects, AABB.h:10-14) lets a NaN box overlap every box.  The
// sweep order is unaffected: such a box's sort key sits at an end of its
// scene, where no other position's scan reads it as a stop.
__global__ void k_bp_gather(Bodies b, const int* sorted, double* sb, int skip_nonfinite) {
	int p = blockIdx.x * blockDim.x + threadIdx.x;
	const int m = b.nl;
	if (p >= m) return;
	int i = sorted[p];
	double v[6];
	bool fin = true;
	for (int c = 0; c < 6; c++) {
		v[c] = b.aabb[c * b.n + i];
		fin = fin && isfinite(v[c]);
	}
	if (skip_nonfinite && !fin)
		for (int c = 0; c < 6; c++) v[c] = c < 3 ? __builtin_inf() : -__builtin_inf();
	for (int c = 0; c < 6; c++) sb[c * m + p] = v[c];
}

// Sweep end of sorted position p: the first q > p of p's scene with
// min.x[q] > max.x[p] (a forward scan stops there; min.x is ascending in the
// scene except for NaNs, which sort to the ends and never stop a scan, so a
// NaN found first means the scan runs to the scene's end).  nchunk[p] = chunks
// of BP_CH positions in (p, ext[p]).
__global__ void k_bp_extent(int n, const int* sorted, const double* sb, const int* scene_of_body, const int* scene_body_off, int* ext,
                            int* nchunk) {
	int p = blockIdx.x * blockDim.x + threadIdx.x;
	if (p >= n) return;
	int i = sorted[p];
	int end = scene_body_off[scene_of_body[i] + 1];
	double amax0 = sb[3 * n + p];
	int lo = p + 1, hi = end;
	while (lo < hi) {
		int mid = (lo + hi) >> 1;
		double v = sb[mid];
		bool stop = v > amax0 || (v != v && !signbit(v));
		if (stop) hi = mid;
		else lo = mid + 1;
	}
	if (lo < end && sb[lo] != sb[lo]) lo = end;
	ext[p] = lo;
	nchunk[p] = (lo - p - 1 + BP_CH - 1) / BP_CH;
}

__global__ void k_bp_owner(int n, const int* choff, int* owner) {
	int p = blockIdx.x * blockDim.x + threadIdx.x;
	if (p >= n) return;
	for (int c = choff[p]; c < choff[p + 1]; c++) owner[c] = p;
}

// One thread per chunk (grid-stride; the chunk count is read on the device).
__global__ void k_bp_sweep(int n, const int* sorted, const double* sb, const int* ext, const int* choff, const int* owner, int* counts,
                           const int* offs, uint64_t* out, int write) {
	const int nch = choff[n];
	for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nch; c += gridDim.x * blockDim.x) {
		int p = owner[c];
		int q0 = p + 1 + (c - choff[p]) * BP_CH;
		int q1 = min(ext[p], q0 + BP_CH);
		int i = sorted[p];
		double amin0 = sb[p], amin1 = sb[n + p], amin2 = sb[2 * n + p];
		double amax0 = sb[3 * n + p], amax1 = sb[4 * n + p], amax2 = sb[5 * n + p];
		int cnt = 0;
		int base = write ? offs[c] : 0;
		for (int q = q0; q < q1; q++) {
			double bmin0 = sb[q], bmin1 = sb[n + q], bmin2 = sb[2 * n + q];
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
		if (!write) counts[c] = cnt;
	}
}

static inline int nblk(int n, int t) { return (n + t - 1) / t; }

void launch_bp_keys(const Bodies& b, uint64_t* keys, int* vals, hipStream_t s) {
	if (b.nl) hipLaunchKernelGGL(k_bp_keys, dim3(nblk(b.nl, 256)), dim3(256), 0, s, b, keys, vals);
}
void launch_bp_gather(const Bodies& b, const int* sorted, double* sb, int skip_nonfinite, hipStream_t s) {
	if (b.nl) hipLaunchKernelGGL(k_bp_gather, dim3(nblk(b.nl, 256)), dim3(256), 0, s, b, sorted, sb, skip_nonfinite);
}
void launch_bp_extent(int n, const int* sorted, const double* sb, const int* scene_of_body, const int* scene_body_off, int* ext,
                      int* nchunk, hipStream_t s) {
	if (n) hipLaunchKernelGGL(k_bp_extent, dim3(nblk(n, 256)), dim3(256), 0, s, n, sorted, sb, scene_of_body, scene_body_off, ext, nchunk);
}
void launch_bp_owner(int n, const int* choff, int* owner, hipStream_t s) {
	if (n) hipLaunchKernelGGL(k_bp_owner, dim3(nblk(n, 256)), dim3(256), 0, s, n, choff, owner);
}
void launch_bp_sweep(int n, int nch, const int* sorted, const double* sb, const int* ext, const int* choff, const int* owner, int* counts,
                     const int* offs, uint64_t* out, int write, hipStream_t s) {
	if (nch) hipLaunchKernelGGL(k_bp_sweep, dim3(std::min(nblk(nch, 256), 4096)), dim3(256), 0, s, n, sorted, sb, ext, choff, owner, counts,
	                            offs, out, write);
}

}  // namespace cpx
