// Contact persistence, constraint-graph edges and edge colouring.
//
// Persistence restates PhysicsEngine::addContact (PhysicsEngine.cpp:1062-1111)
// and the contact life cycle (:1378-1446) over a cache kept sorted by
// (pair key, insertion order) — the per-edge vectors of the reference laid end
// to end.  A step's new manifold points (sorted by pair key) are matched to the
// cache by binary search on the key and a MagicID compare; matched contacts are
// refreshed in place, unmatched ones are merged in after the existing contacts
// of their pair.  The merge needs no sort: each element's output position is
// its rank in its own list plus a binary-search count in the other list.
//
// Removal of expired contacts (memory_life < 0 at the first substep, neither
// body asleep) is deferred to the next step's merge: the reference removes
// them before any later read, so dropping entries whose life fell below -1
// after the end-of-step decrement (contact_removed) yields the same cache.
//
// Colouring: one thread per scene greedily colours the scene's edges (body
// pairs with at least one live contact) so that no two edges of a colour share
// a dynamic body; the PGS kernel then sweeps colour by colour.
//
// Roofline: HBM-bound gathers/scatters; per contact ~240 B read + written per
// step (cache fields), per new point ~100 B.
#include "cp_internal.h"

namespace cpx {

__global__ void k_np_gather(Manifolds M, int mcap, const uint64_t* pairs, int npairs, const int* off, NewPts np, int cap) {
	int s = blockIdx.x * blockDim.x + threadIdx.x;
	if (s >= npairs) return;
	int c = M.count[s];
	int o = off[s];
	for (int k = 0; k < c; k++) {
		int j = o + k;
		np.key[j] = pairs[s];
		np.cid[j] = M.magic[(size_t)s * 8 + 2 * k];
		np.bid[j] = M.magic[(size_t)s * 8 + 2 * k + 1];
		for (int d = 0; d < 3; d++) {
			np.pos[d * cap + j] = M.pts[(size_t)s * 16 + 4 * k + d];
			np.nrm[d * cap + j] = M.normal[d * mcap + s];
			np.mat[d * cap + j] = M.mat[(size_t)s * 12 + 3 * k + d];
		}
		np.pen[j] = M.pts[(size_t)s * 16 + 4 * k + 3];
	}
}

__device__ __forceinline__ int lower_bound_u64(const uint64_t* a, int n, uint64_t v) {
	int lo = 0, hi = n;
	while (lo < hi) {
		int m = (lo + hi) >> 1;
		if (a[m] < v) lo = m + 1;
		else hi = m;
	}
	return lo;
}
__device__ __forceinline__ int upper_bound_u64(const uint64_t* a, int n, uint64_t v) {
	int lo = 0, hi = n;
	while (lo < hi) {
		int m = (lo + hi) >> 1;
		if (a[m] <= v) lo = m + 1;
		else hi = m;
	}
	return lo;
}

// Local frames of a new point (PhysicsEngine.cpp:1067-1071).
__device__ void local_frames(const Bodies& b, int i1, int i2, V3 p, V3 n, V3* lp1, V3* lp2, V3* ln1) {
	const int N = b.n;
	Q q1{b.q[i1], b.q[N + i1], b.q[2 * N + i1], b.q[3 * N + i1]};
	Q q2{b.q[i2], b.q[N + i2], b.q[2 * N + i2], b.q[3 * N + i2]};
	V3 c1{b.com[i1], b.com[N + i1], b.com[2 * N + i1]};
	V3 c2{b.com[i2], b.com[N + i2], b.com[2 * N + i2]};
	M3 m1 = q_rot(q_conj(q1)), m2 = q_rot(q_conj(q2));
	*ln1 = m1 * n;
	*lp1 = m1 * (p - c1);
	*lp2 = m2 * (p - c2);
}

// Match each new point against the cache; refresh matched contacts in place.
// Within one pair, repeated MagicIDs behave as sequential addContact calls:
// the first occurrence owns the slot, the last occurrence's data is kept.
__global__ void k_np_match(NewPts np, int nnew, int cap, Contacts C, int nold, Bodies b, Params P) {
	int j = blockIdx.x * blockDim.x + threadIdx.x;
	if (j >= nnew) return;
	uint64_t key = np.key[j], cid = np.cid[j], bid = np.bid[j];
	bool first = true, last = true;
	int lastj = j;
	for (int k = j - 1; k >= 0 && np.key[k] == key; k--)
		if (np.cid[k] == cid && np.bid[k] == bid) first = false;
	for (int k = j + 1; k < nnew && np.key[k] == key; k++)
		if (np.cid[k] == cid && np.bid[k] == bid) { last = false; lastj = k; }
	int m = -1;
	for (int k = lower_bound_u64(C.key, nold, key); k < nold && C.key[k] == key; k++) {
		if (contact_removed(C, k, b.asleep_rm)) continue;  // removed at the previous step's first substep
		if (C.cid[k] == cid && C.bid[k] == bid) { m = k; break; }
	}
	np.is_new[j] = (m < 0 && first) ? 1 : 0;
	np.src[j] = lastj;
	if (m >= 0 && last) {
		int i1 = (int)(key >> 32), i2 = (int)(key & 0xFFFFFFFFu);
		V3 p{np.pos[j], np.pos[cap + j], np.pos[2 * cap + j]};
		V3 n{np.nrm[j], np.nrm[cap + j], np.nrm[2 * cap + j]};
		V3 lp1, lp2, ln1;
		local_frames(b, i1, i2, p, n, &lp1, &lp2, &ln1);
		int K = C.cap;
		C.mu[m] = C.sready[m] ? np.mat[2 * cap + j] : np.mat[cap + j];
		C.pen[m] = np.pen[j];
		C.ln1[m] = ln1.x; C.ln1[K + m] = ln1.y; C.ln1[2 * K + m] = ln1.z;
		C.lp1[m] = lp1.x; C.lp1[K + m] = lp1.y; C.lp1[2 * K + m] = lp1.z;
		C.lp2[m] = lp2.x; C.lp2[K + m] = lp2.y; C.lp2[2 * K + m] = lp2.z;
		C.life[m] = P.contact_life;
		C.live[m] = 1;
	}
}

__global__ void k_old_keep(const Contacts C, int nold, int* keep, const int* prot) {
	int k = blockIdx.x * blockDim.x + threadIdx.x;
	if (k < nold) keep[k] = contact_removed(C, k, prot) ? 0 : 1;
	if (k == nold) keep[k] = 0;
}

__global__ void k_new_keys(NewPts np, int nnew, const int* rank, uint64_t* nk) {
	int j = blockIdx.x * blockDim.x + threadIdx.x;
	if (j < nnew && np.is_new[j]) nk[rank[j]] = np.key[j];
}

__device__ void copy_contact(const Contacts& S, int i, Contacts& D, int o) {
	int KS = S.cap, KD = D.cap;
	D.key[o] = S.key[i]; D.cid[o] = S.cid[i]; D.bid[o] = S.bid[i]; D.seq[o] = S.seq[i];
	for (int d = 0; d < 3; d++) {
		D.lp1[d * KD + o] = S.lp1[d * KS + i];
		D.lp2[d * KD + o] = S.lp2[d * KS + i];
		D.ln1[d * KD + o] = S.ln1[d * KS + i];
		D.fu[d * KD + o] = S.fu[d * KS + i];
		D.fw[d * KD + o] = S.fw[d * KS + i];
	}
	D.pen[o] = S.pen[i]; D.bounce[o] = S.bounce[i]; D.mu[o] = S.mu[i];
	D.life[o] = S.life[i]; D.live[o] = S.live[i]; D.sready[o] = S.sready[i];
	D.imp_n[o] = S.imp_n[i];
	D.imp_t[o] = S.imp_t[i];
	D.imp_t[KD + o] = S.imp_t[KS + i];
}

__global__ void k_merge_old(Contacts S, int nold, const int* keep_scan, const uint64_t* nk, int nnewc, Contacts D, const int* prot) {
	int k = blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= nold || contact_removed(S, k, prot)) return;
	int o = keep_scan[k] + lower_bound_u64(nk, nnewc, S.key[k]);
	copy_contact(S, k, D, o);
}

// New contact (PhysicsEngine.cpp:1096-1110 / PersistentConstraint.h:51-56).
__global__ void k_merge_new(NewPts np, int nnew, int cap, const int* rank, Contacts S, int nold, const int* keep_scan, Contacts D,
                            Bodies b, Params P, uint64_t seq0) {
	int j = blockIdx.x * blockDim.x + threadIdx.x;
	if (j >= nnew || !np.is_new[j]) return;
	uint64_t key = np.key[j];
	int u = upper_bound_u64(S.key, nold, key);
	int o = rank[j] + keep_scan[u];
	int s = np.src[j];
	int i1 = (int)(key >> 32), i2 = (int)(key & 0xFFFFFFFFu);
	V3 p{np.pos[s], np.pos[cap + s], np.pos[2 * cap + s]};
	V3 n{np.nrm[s], np.nrm[cap + s], np.nrm[2 * cap + s]};
	V3 lp1, lp2, ln1;
	local_frames(b, i1, i2, p, n, &lp1, &lp2, &ln1);
	int K = D.cap;
	D.key[o] = key; D.cid[o] = np.cid[j]; D.bid[o] = np.bid[j]; D.seq[o] = seq0 + (uint64_t)j;
	D.lp1[o] = lp1.x; D.lp1[K + o] = lp1.y; D.lp1[2 * K + o] = lp1.z;
	D.lp2[o] = lp2.x; D.lp2[K + o] = lp2.y; D.lp2[2 * K + o] = lp2.z;
	D.ln1[o] = ln1.x; D.ln1[K + o] = ln1.y; D.ln1[2 * K + o] = ln1.z;
	for (int d = 0; d < 3; d++) { D.fu[d * K + o] = 0; D.fw[d * K + o] = 0; }
	D.pen[o] = np.pen[s];
	D.bounce[o] = np.mat[s];
	D.mu[o] = np.mat[cap + s];  // new contacts take the kinetic coefficient
	D.life[o] = P.contact_life;
	D.live[o] = 1;
	D.sready[o] = 0;
	D.imp_n[o] = 0;
	D.imp_t[o] = 0;
	D.imp_t[K + o] = 0;
}

// Last substep (PhysicsEngine.cpp:1387-1392): memory_life--, is_live = false.
__global__ void k_age(Contacts C, int nc, const int* ncp) {
	const int n = ncp ? *ncp : nc;
	for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
		C.life[k] -= 1;
		C.live[k] = 0;
	}
}

__global__ void k_set_count(int* p, int v) {
	if (threadIdx.x == 0) *p = v;
}

// ------------------------------------------------------------- edges
__global__ void k_edge_heads(const Contacts C, int n, int* head) {
	int k = blockIdx.x * blockDim.x + threadIdx.x;
	if (k < n) head[k] = (k == 0 || C.key[k] != C.key[k - 1]) ? 1 : 0;
}
__global__ void k_edge_build(const Contacts C, int n, const int* head, const int* eidx, uint64_t* ekey, int* efirst, int* ecnt,
                             int* elive) {
	int k = blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= n || !head[k]) return;
	int e = eidx[k];
	int c = 0, live = 0;
	for (int t = k; t < n && C.key[t] == C.key[k]; t++) { c++; live |= C.live[t]; }
	ekey[e] = C.key[k];
	efirst[e] = k;
	ecnt[e] = c;
	elive[e] = live;
}

#define MAXC CP_MAXC

// Colour sets of a scene's bodies: two 64-bit words per body (colours 0-63
// and 64-127), word w of body i at w{0,1}[i].  Word 0 decides almost every
// edge (word 1 is read only when a body already holds 64 colours), so it
// lives in LDS for scenes of up to CP_COLOR_LDS_WORDS bodies; word 1 lives in
// LDS too when both fit, else in global memory.
#define CP_COLOR_LDS_WORDS 20352  // 159 KB of LDS (with the kernel's own colour counts, within the 160 KB limit)
struct ColorSets {
	uint64_t* w0;
	uint64_t* w1;
	bool g0, g1;  // the word is in global memory
};

// Greedy colouring of one scene's edges [e0, e1) in order, 64 at a time:
// each lane loads one edge (scene-local bodies, live and dynamic flags), then
// the wave walks the 64 edges in order with the edge's fields read from its
// lane (readlane, so they are wave-uniform) and gives each the lowest colour
// free at its dynamic bodies (colour sets, updated by lane 0; an update to a
// word in global memory is fenced before the next edge reads it).  The rank of
// an edge within its colour (edges of that colour before it) is then found for
// all 64 at once: lanes of equal colour are matched bit by bit with ballots
// and counted below, on top of the running count cnt[c].  ecolor = colour |
// rank << 8, -1 for an edge with no live contact.  Returns the largest colour
// used.
__device__ int color_edges(const Bodies& b, int b0, int e0, int e1, const uint64_t* ekey, const int* elive, const ColorSets& M, int* cnt,
                           int* ecolor, int lane) {
	const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
	int maxc = -1;
	for (int base = e0; base < e1; base += 64) {
		const int e = base + lane, n = min(64, e1 - base);
		int lo = 0, hi = 0, fl = 0;
		if (e < e1) {
			uint64_t k = ekey[e];
			int glo = (int)(k >> 32), ghi = (int)(k & 0xFFFFFFFFu);
			fl = (elive[e] ? 1 : 0) | (b.movement[glo] == 0 ? 2 : 0) | (b.movement[ghi] == 0 ? 4 : 0);
			lo = glo - b0;
			hi = ghi - b0;
		}
		int myc = -1;
		for (int j = 0; j < n; j++) {
			const int jfl = __builtin_amdgcn_readlane(fl, j);
			if (!(jfl & 1)) continue;
			const int jlo = __builtin_amdgcn_readlane(lo, j), jhi = __builtin_amdgcn_readlane(hi, j);
			const bool dl = (jfl & 2) != 0, dh = (jfl & 4) != 0;
			uint64_t m0 = (dl ? M.w0[jlo] : 0) | (dh ? M.w0[jhi] : 0);
			int c;
			if (~m0) c = __builtin_ctzll(~m0);
			else {
				uint64_t m1 = (dl ? M.w1[jlo] : 0) | (dh ? M.w1[jhi] : 0);
				c = (~m1) ? 64 + __builtin_ctzll(~m1) : MAXC;
			}
			if (c < MAXC) {
				uint64_t* w = c < 64 ? M.w0 : M.w1;
				if (lane == 0) {
					uint64_t bit = 1ull << (c & 63);
					if (dl) w[jlo] |= bit;
					if (dh) w[jhi] |= bit;
				}
				if (c < 64 ? M.g0 : M.g1) {  // wave-uniform: the next edge's loads see lane 0's stores
					__builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
					__builtin_amdgcn_wave_barrier();
				}
			}
			if (lane == j) myc = c;
			maxc = c > maxc ? c : maxc;
		}
		// lanes with the same colour (colours 0..MAXC fit in 8 bits; -1 has
		// all 8 set and is never ranked)
		uint64_t peers = ~0ull;
#pragma unroll
		for (int bit = 0; bit < 8; bit++) {
			const bool set = (myc >> bit) & 1;
			const uint64_t m = __ballot(set);
			peers &= set ? m : ~m;
		}
		int rank = 0;
		if (myc >= 0) rank = cnt[myc] + __popcll(peers & below);
		__builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
		__builtin_amdgcn_wave_barrier();
		if (myc >= 0 && (peers & ~below & ~(1ull << lane)) == 0) cnt[myc] += __popcll(peers);
		__builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
		__builtin_amdgcn_wave_barrier();
		if (e < e1) ecolor[e] = myc < 0 ? -1 : (myc | (rank << 8));
	}
	return maxc;
}

// One wave per scene (scene_list[blockIdx.x], or blockIdx.x without a list):
// greedy edge colouring in edge (cache key) order (color_edges), then a
// counting sort of the scene's live edges by colour.
// lds_words: 64-bit words of LDS per workgroup (at most CP_COLOR_LDS_WORDS).
// A scene keeps both colour-set words in LDS when 2 x its bodies fit, word 0
// when its bodies fit, else both words in b.color_mask.  Colour MAXC collects
// edges that found no free colour; the solver runs that bucket on a single
// thread.
__global__ __launch_bounds__(64) void k_color(Bodies b, const int* scene_body_off, const int* scene_list, const uint64_t* ekey,
                                              const int* elive, int nedges, int* ecolor, int* order, int* color_off, int* ncolors,
                                              int* edge_range, int lds_words) {
	extern __shared__ uint64_t smask_lds[];
	__shared__ int cnt[MAXC + 1];
	const int s = scene_list ? scene_list[blockIdx.x] : blockIdx.x, lane = threadIdx.x;
	const int b0 = scene_body_off[s], b1 = scene_body_off[s + 1], nbs = b1 - b0;
	const int e0 = lower_bound_u64(ekey, nedges, (uint64_t)b0 << 32), e1 = lower_bound_u64(ekey, nedges, (uint64_t)b1 << 32);
	uint64_t* gmask = reinterpret_cast<uint64_t*>(b.color_mask) + 2 * (size_t)b0;  // [2][nbs] words of this scene
	ColorSets M;
	M.g0 = nbs > lds_words;
	M.g1 = 2 * nbs > lds_words;
	M.w0 = M.g0 ? gmask : smask_lds;
	M.w1 = M.g1 ? gmask + nbs : smask_lds + nbs;
	for (int i = lane; i < nbs; i += 64) {
		if (!M.g0) M.w0[i] = 0;
		if (!M.g1) M.w1[i] = 0;
	}
	if (M.g1)
		for (int e = e0 + lane; e < e1; e += 64) {  // only bodies on an edge need a clear colour set
			int lo = (int)(ekey[e] >> 32) - b0, hi = (int)(ekey[e] & 0xFFFFFFFFu) - b0;
			M.w1[lo] = 0;
			M.w1[hi] = 0;
			if (M.g0) {
				M.w0[lo] = 0;
				M.w0[hi] = 0;
			}
		}
	for (int c = lane; c <= MAXC; c += 64) cnt[c] = 0;
	if (lane == 0) {
		edge_range[2 * s] = e0;
		edge_range[2 * s + 1] = e1;
	}
	__syncthreads();
	const int maxc = color_edges(b, b0, e0, e1, ekey, elive, M, cnt, ecolor, lane);
	__syncthreads();
	int* co = color_off + (size_t)s * (MAXC + 2);
	if (lane == 0) {
		int acc = 0;
		for (int c = 0; c <= MAXC; c++) {
			co[c] = acc;
			int v = cnt[c];
			cnt[c] = acc;
			acc += v;
		}
		co[MAXC + 1] = acc;
		ncolors[s] = maxc + 1;
	}
	__syncthreads();
	for (int e = e0 + lane; e < e1; e += 64) {
		int v = ecolor[e];
		if (v < 0) continue;
		order[e0 + cnt[v & 0xFF] + (v >> 8)] = e;
	}
}

// Descriptor of each colour-sorted edge slot (order[x] = -1: unused slot).
__global__ void k_edge_slots(Bodies b, Contacts C, const int* order, int nslots, const int* efirst, const int* ecnt, int4* desc,
                             uint32_t* lmask) {
	int x = blockIdx.x * blockDim.x + threadIdx.x;
	if (x >= nslots) return;
	int e = order[x];
	if (e < 0) return;
	int k0 = efirst[e], cnt = ecnt[e];
	int i1 = (int)(C.key[k0] >> 32), i2 = (int)(C.key[k0] & 0xFFFFFFFFu);
	uint32_t m = 0;
	for (int j = 0; j < cnt && j < 32; j++)
		if (C.live[k0 + j]) m |= 1u << j;
	uint32_t y = (uint32_t)(cnt & 0xFFFF) | (cnt > 32 ? (uint32_t)ED_LONG : 0u) | (b.movement[i1] == 0 ? (uint32_t)ED_DYN_A : 0u) |
	             (b.movement[i2] == 0 ? ED_DYN_B : 0u);
	desc[x] = make_int4(k0, (int)y, i1, i2);
	lmask[x] = m;
}

static inline int nblk(int n, int t) { return (n + t - 1) / t; }

void launch_edge_slots(const Bodies& b, const Contacts& C, const int* order, int nslots, const int* efirst, const int* ecnt, int4* desc,
                       uint32_t* lmask, hipStream_t s) {
	if (nslots) hipLaunchKernelGGL(k_edge_slots, dim3(nblk(nslots, 256)), dim3(256), 0, s, b, C, order, nslots, efirst, ecnt, desc, lmask);
}

void launch_np_gather(const Manifolds& M, int mcap, const uint64_t* pairs, int npairs, const int* off, const NewPts& np, int cap,
                      hipStream_t s) {
	if (npairs) hipLaunchKernelGGL(k_np_gather, dim3(nblk(npairs, 256)), dim3(256), 0, s, M, mcap, pairs, npairs, off, np, cap);
}
void launch_np_match(const NewPts& np, int nnew, int cap, const Contacts& C, int nold, const Bodies& b, const Params& P,
                     hipStream_t s) {
	if (nnew) hipLaunchKernelGGL(k_np_match, dim3(nblk(nnew, 128)), dim3(128), 0, s, np, nnew, cap, C, nold, b, P);
}
void launch_old_keep(const Contacts& C, int nold, int* keep, const int* prot, hipStream_t s) {
	hipLaunchKernelGGL(k_old_keep, dim3(nblk(nold + 1, 256)), dim3(256), 0, s, C, nold, keep, prot);
}
void launch_new_keys(const NewPts& np, int nnew, const int* rank, uint64_t* nk, hipStream_t s) {
	if (nnew) hipLaunchKernelGGL(k_new_keys, dim3(nblk(nnew, 256)), dim3(256), 0, s, np, nnew, rank, nk);
}
void launch_merge(const Contacts& S, int nold, const int* keep_scan, const uint64_t* nk, int nnewc, const NewPts& np, int nnew, int cap,
                  const int* rank, const Contacts& D, const Bodies& b, const Params& P, uint64_t seq0, hipStream_t s) {
	if (nold) hipLaunchKernelGGL(k_merge_old, dim3(nblk(nold, 128)), dim3(128), 0, s, S, nold, keep_scan, nk, nnewc, D, b.asleep_rm);
	if (nnew)
		hipLaunchKernelGGL(k_merge_new, dim3(nblk(nnew, 128)), dim3(128), 0, s, np, nnew, cap, rank, S, nold, keep_scan, D, b, P, seq0);
}
void launch_age(const Contacts& C, int n, hipStream_t s, const int* ncp) {
	if (!n) return;
	int g = ncp ? std::max(1, std::min(nblk(n, 256), 4096)) : nblk(n, 256);
	hipLaunchKernelGGL(k_age, dim3(g), dim3(256), 0, s, C, n, ncp);
}
void launch_set_count(int* p, int v, hipStream_t s) { hipLaunchKernelGGL(k_set_count, dim3(1), dim3(64), 0, s, p, v); }
void launch_edge_heads(const Contacts& C, int n, int* head, hipStream_t s) {
	if (n) hipLaunchKernelGGL(k_edge_heads, dim3(nblk(n, 256)), dim3(256), 0, s, C, n, head);
}
void launch_edge_build(const Contacts& C, int n, const int* head, const int* eidx, uint64_t* ekey, int* efirst, int* ecnt, int* elive,
                       hipStream_t s) {
	if (n) hipLaunchKernelGGL(k_edge_build, dim3(nblk(n, 256)), dim3(256), 0, s, C, n, head, eidx, ekey, efirst, ecnt, elive);
}
__global__ void k_kill_contacts(Contacts C, int n, const int* ranges, int nranges) {
	int k = blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= n) return;
	int lo = (int)(C.key[k] >> 32);
	for (int r = 0; r < nranges; r++)
		if (lo >= ranges[2 * r] && lo < ranges[2 * r] + ranges[2 * r + 1]) {
			C.life[k] = -100;
			C.live[k] = 0;
			return;
		}
}
void launch_kill_contacts(const Contacts& C, int n, const int* ranges, int nranges, hipStream_t s) {
	if (n > 0 && nranges > 0) hipLaunchKernelGGL(k_kill_contacts, dim3(nblk(n, 256)), dim3(256), 0, s, C, n, ranges, nranges);
}

void launch_color(const Bodies& b, const int* scene_body_off, int nscenes, const uint64_t* ekey, const int* elive, int nedges,
                  int* ecolor, int* order, int* color_off, int* ncolors, int* edge_range, int max_bodies, hipStream_t s,
                  const int* scene_list) {
	if (!nscenes) return;
	// colour-set words in LDS: both words of every scene when the largest fits,
	// else word 0 of scenes up to CP_COLOR_LDS_WORDS bodies (159 KB).  The LDS a
	// workgroup takes bounds how many run on a CU, so callers launch scenes of
	// similar size together.
	int lds_words = std::min(2 * max_bodies, CP_COLOR_LDS_WORDS);
	static bool attr = false;
	if (!attr) {
		if (hipFuncSetAttribute((const void*)k_color, hipFuncAttributeMaxDynamicSharedMemorySize, 8 * CP_COLOR_LDS_WORDS) != hipSuccess)
			(void)hipGetLastError();  // the launch then reports the error itself
		attr = true;
	}
	hipLaunchKernelGGL(k_color, dim3(nscenes), dim3(64), (size_t)8 * lds_words, s, b, scene_body_off, scene_list, ekey, elive, nedges, ecolor,
	                   order, color_off, ncolors, edge_range, lds_words);
}

}  // namespace cpx
