// Island search on the device: PhysicsEngine::sleepOrSolveIslands
// (PhysicsEngine.cpp:1536-1613) with bfsVisitAll (:1263-1283) over the
// post-narrowphase contact cache and the joints, and the island sleep
// decision (:1598-1606, readyToSleep :1256-1258).
//
// The constraint graph's edges are body pairs holding cached contacts or
// joints; a node lists its edges by the other body's id
// (ConstraintGraphNode::insertNewEdge, :1984-1995).  Islands are the
// connected components over dynamic bodies (fixed bodies are leaves and
// never bridge two islands), taken in the order of their first body in the
// engine's node map; each island's rows come from a depth-first walk from
// that body in the reference's order.
//
// Stages (one launch each, no host round trip until the island count):
//   merge     contact pairs and joint pairs into one key-sorted edge table
//   adjacency (body, neighbour) entries sorted by key, CSR offsets per body
//   components union-find hooking the root of later node-map rank under the
//             earlier one, so every component's root is its first body
//   counts    per component: bodies, adjacency entries, rows, edges, joints,
//             and whether every body is ready to sleep
//   tables    an exclusive scan over node-map positions numbers the islands
//             that are solved this step and places their ranges
//   walk      one thread per island: the depth-first walk, the island's row
//             list, and for coloured islands the greedy edge colouring,
//             island-local body indices and their maps
#include "cp_internal.h"

namespace cpx {

namespace {
__device__ __forceinline__ int ld_relaxed(const int* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ __forceinline__ void st_relaxed(int* p, int v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

// Root of x with path halving; parents only ever move to bodies of earlier
// node-map rank, so concurrent halving and hooking keep every path valid.
__device__ int uf_find(int* parent, int x) {
	while (true) {
		int p = ld_relaxed(parent + x);
		if (p == x) return x;
		int g = ld_relaxed(parent + p);
		if (g != p) st_relaxed(parent + x, g);
		x = p;
	}
}
__device__ void uf_union(int* parent, const int* rank, int a, int b) {
	while (true) {
		a = uf_find(parent, a);
		b = uf_find(parent, b);
		if (a == b) return;
		if (rank[a] < rank[b]) {
			int t = a;
			a = b;
			b = t;
		}
		// a has the later rank: hook it under b
		int old = atomicCAS(parent + a, a, b);
		if (old == a) return;
	}
}
}  // namespace

// Sort keys of the two edge sources: contact pairs (values 0..ne-1) and
// joint pair groups (values ne..ne+ng-1).
__global__ void k_ib_keys(const uint64_t* ekey, int ne, const uint64_t* jgkey, int ng, uint64_t* keys, int* vals) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= ne + ng) return;
	keys[i] = i < ne ? ekey[i] : jgkey[i - ne];
	vals[i] = i;
}

__global__ void k_ib_heads(const uint64_t* skeys, int n, int* head) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i > n) return;
	head[i] = i < n && (i == 0 || skeys[i] != skeys[i - 1]) ? 1 : 0;
}

// Merged edge e = midx[i] of head entry i: lo, hi, first contact, contacts,
// live contacts, joints jord[j0, j1), active joints among them (an
// inequality joint adds its row only while active, JointAct); and its two
// adjacency entries.  Slots past the merged count get adjacency keys that
// sort last.
__global__ void k_ib_edges(const uint64_t* skeys, const int* svals, int n, const int* head, const int* midx, int ne, const int* efirst,
                           const int* ecnt, const int* live, const int* jgoff, const int* jord, JointAct jact, int* me,
                           uint64_t* adjk, int* adjv) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	const int M = midx[n];
	if (i >= M) {
		adjk[2 * i] = ~0ull;
		adjk[2 * i + 1] = ~0ull;
		adjv[2 * i] = -1;
		adjv[2 * i + 1] = -1;
	}
	if (!head[i]) return;
	const int e = midx[i];
	const uint64_t K = skeys[i];
	int E[8] = {(int)(K >> 32), (int)(K & 0xFFFFFFFFu), -1, 0, 0, 0, 0, 0};
	for (int x = i; x < n && skeys[x] == K; x++) {
		int v = svals[x];
		if (v < ne) {
			E[2] = efirst[v];
			E[3] = ecnt[v];
			int nl = 0;
			for (int k = E[2]; k < E[2] + E[3]; k++) nl += live[k] != 0;
			E[4] = nl;
		} else {
			E[5] = jgoff[v - ne];
			E[6] = jgoff[v - ne + 1];
			for (int x2 = E[5]; x2 < E[6]; x2++) E[7] += jact(jord[x2]);
		}
	}
	int4* o = reinterpret_cast<int4*>(me + 8 * (size_t)e);
	o[0] = int4{E[0], E[1], E[2], E[3]};
	o[1] = int4{E[4], E[5], E[6], E[7]};
	adjk[2 * e] = (uint64_t)(uint32_t)E[0] << 32 | (uint32_t)E[1];
	adjk[2 * e + 1] = (uint64_t)(uint32_t)E[1] << 32 | (uint32_t)E[0];
	adjv[2 * e] = e;
	adjv[2 * e + 1] = e;
}

// CSR offsets of the sorted adjacency: aoff[b] = first entry of body b.
__global__ void k_ib_aoff(const uint64_t* adjk, int n2, int nb, int* aoff) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i > n2) return;
	auto body = [&](int x) { return x >= n2 ? nb : (int)min<uint64_t>(adjk[x] >> 32, (uint64_t)nb); };
	int b = body(i), pb = i == 0 ? -1 : body(i - 1);
	for (int x = pb + 1; x <= b && x <= nb; x++) aoff[x] = i;
}

__global__ void k_ib_init(int* parent, IslCnt* acc, int nb) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= nb) return;
	parent[i] = i;
	acc[i] = IslCnt{0, 0, 0, 0, 0, 0, 0};
}

__global__ void k_ib_union(const int* me, const int* midx, int n, const int* movement, const int* rank, int* parent) {
	int e = blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= n || e >= midx[n]) return;
	int lo = me[8 * (size_t)e], hi = me[8 * (size_t)e + 1];
	if (movement[lo] == 0 && movement[hi] == 0) uf_union(parent, rank, lo, hi);
}

// readyToSleep (PhysicsEngine.cpp:1256-1258) of dynamic body i.
__device__ __forceinline__ bool ready_to_sleep(const SleepDev& S, int i) { return S.counter[i] >= S.delay; }

// Component of every body (-1 for non-dynamic ones) and its body counts.
__global__ void k_ib_bodies(int nb, const int* movement, const int* aoff, int* parent, int* comp, IslCnt* acc, SleepDev S) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= nb) return;
	if (movement[i] != 0) {
		comp[i] = -1;
		return;
	}
	int r = uf_find(parent, i);
	comp[i] = r;
	atomicAdd(&acc[r].bodies, 1);
	atomicAdd(&acc[r].adj, aoff[i + 1] - aoff[i]);
	if (!S.sleeping || !ready_to_sleep(S, i)) atomicOr(&acc[r].notready, 1);
}

// The rows an edge gives its island: none when neither end is an awake
// dynamic body (bfsVisitAll skips an asleep node's edges to asleep or fixed
// neighbours); else its active joints and two rows per live contact.
__device__ __forceinline__ bool edge_emitted(const int* E, const int* movement, const int* asleep) {
	if (E[7] == 0 && E[4] == 0) return false;
	bool a = movement[E[0]] == 0 && !(asleep && asleep[E[0]]);
	bool b = movement[E[1]] == 0 && !(asleep && asleep[E[1]]);
	return a || b;
}

__global__ void k_ib_edge_counts(const int* me, const int* midx, int n, const int* movement, const int* comp, const int* asleep, IslCnt* acc) {
	int e = blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= n || e >= midx[n]) return;
	const int* E = me + 8 * (size_t)e;
	int d = movement[E[0]] == 0 ? E[0] : (movement[E[1]] == 0 ? E[1] : -1);
	if (d < 0 || !edge_emitted(E, movement, asleep)) return;
	int r = comp[d];
	int nj = E[7];
	atomicAdd(&acc[r].rows, nj + 2 * E[4]);
	atomicAdd(&acc[r].edges, 1);
	atomicAdd(&acc[r].joints, nj);
}

// Scan input at node-map position p: the counts of the component rooted at
// that body when the island is solved this step (it has rows and does not
// go to sleep), else zero; position nb is zero (the scan's totals).
__global__ void k_ib_flags(int nb, const int* rord, const int* comp, const IslCnt* acc, int sleeping, IslCnt* in) {
	int p = blockIdx.x * blockDim.x + threadIdx.x;
	if (p > nb) return;
	IslCnt z{0, 0, 0, 0, 0, 0, 0};
	if (p == nb) {
		in[p] = z;
		return;
	}
	int b = rord[p];
	if (comp[b] != b) {
		in[p] = z;
		return;
	}
	IslCnt c = acc[b];
	bool sleeps = sleeping && !c.notready;
	if (c.rows == 0 || sleeps) {
		in[p] = z;
		return;
	}
	c.n = 1;
	c.notready = 0;
	in[p] = c;
}

// Island tables from the scan; info = {islands, largest island's bodies}.
__global__ void k_ib_tables(int nb, const int* rord, const IslCnt* in, const IslCnt* out, IslTables T, int* info) {
	int p = blockIdx.x * blockDim.x + threadIdx.x;
	if (p > nb) return;
	const IslCnt o = out[p];
	if (p == nb) {
		info[0] = o.n;
		T.off[o.n] = o.rows;
		T.edge_off[o.n] = o.edges;
		T.body_off[o.n] = o.bodies;
		T.stack_off[o.n] = o.adj + o.n;
		T.joint_off[o.n] = o.joints;
		return;
	}
	if (!in[p].n) return;
	const int t = o.n;
	T.root[t] = rord[p];
	T.off[t] = o.rows;
	T.edge_off[t] = o.edges;
	T.body_off[t] = o.bodies;
	T.body_cnt[t] = in[p].bodies;
	T.stack_off[t] = o.adj + t;
	T.joint_off[t] = o.joints;
	atomicMax(info + 1, in[p].bodies);
}

// Islands that go to sleep (sleepOrSolveIslands :1602-1606, RigidBody::sleep
// RigidBody.cpp:322-329): every body of a component whose bodies are all
// ready.
__global__ void k_ib_sleep(int nb, const int* comp, const IslCnt* acc, SleepDev S, double* vel, double* angv) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= nb || comp[i] < 0 || !S.sleeping) return;
	if (acc[comp[i]].notready) return;
	S.asleep[i] = 1;
	for (int c = 0; c < 3; c++) {
		vel[c * nb + i] = 0;
		angv[c * nb + i] = 0;
	}
}

// The walk of island t.  Reference order: rows edge by edge in the walk's
// order (each edge's joints in creation order, then its live contacts in
// cache order, normal row then friction row).  Coloured: the walk's edges
// coloured greedily (lowest colour free at both dynamic ends, CP_MAXC = the
// serial bucket), then emitted colour by colour, with the island's dynamic
// bodies numbered in order of first appearance for the LDS solve.
template <bool COLORED>
__global__ void k_ib_walk(IslWalk W, int nisl) {
	const int t = blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= nisl) return;
	const IslTables& T = W.T;
	const int* mv = W.movement;
	const int* asleep = W.asleep;
	// every range is bounded by the island's counts (stack: one entry per
	// adjacency entry of its bodies, plus the root)
	int* stk = W.stack + T.stack_off[t];
	int* vl = W.vlist + T.body_off[t];
	const int ebase = T.edge_off[t];
	int* el = W.elist + ebase;
	const int scap = T.stack_off[t + 1] - T.stack_off[t], vcap = T.body_off[t + 1] - T.body_off[t], ecap = T.edge_off[t + 1] - ebase;
	int sp = 0, nv = 0, ne = 0;
	stk[sp++] = T.root[t];
	while (sp) {
		int c = stk[--sp];
		if (W.vis[c] || nv == vcap) continue;
		W.vis[c] = 1;
		vl[nv++] = c;
		const bool ca = asleep && asleep[c];
		const int a0 = W.aoff[c], a1 = W.aoff[c + 1];
		for (int x = a0; x < a1; x++) {
			int o = (int)(W.adjk[x] & 0xFFFFFFFFu);
			if (W.vis[o]) continue;
			if (ca && (mv[o] != 0 || (asleep && asleep[o]))) continue;
			const int* E = W.me + 8 * (size_t)W.adjv[x];
			if ((E[7] > 0 || E[4] > 0) && ne < ecap) el[ne++] = W.adjv[x];
		}
		for (int x = a0; x < a1; x++) {
			int o = (int)(W.adjk[x] & 0xFFFFFFFFu);
			if (mv[o] == 0 && !W.vis[o] && sp < scap) stk[sp++] = o;
		}
	}
	for (int x = 0; x < nv; x++) W.vis[vl[x]] = 0;

	int* rows = W.rows + 2 * (size_t)T.off[t];
	const int rcap = T.off[t + 1] - T.off[t];
	int nr = 0;
	auto emit = [&](int e) {
		const int* E = W.me + 8 * (size_t)e;
		for (int x = E[5]; x < E[6] && nr < rcap; x++) {
			if (!W.jact(W.jord[x])) continue;
			rows[2 * nr] = W.jord[x];
			rows[2 * nr + 1] = 2;
			nr++;
		}
		if (E[4] == 0) return;
		for (int k = E[2]; k < E[2] + E[3] && nr + 2 <= rcap; k++) {
			if (!W.live[k]) continue;
			rows[2 * nr] = k; rows[2 * nr + 1] = 0;
			rows[2 * nr + 2] = k; rows[2 * nr + 3] = 1;
			nr += 2;
		}
	};
	if (!COLORED) {
		for (int x = 0; x < ne; x++) emit(el[x]);
	} else {
		int* col = W.ecol + ebase;
		int ncol = 0;
		bool any_serial = false;
		for (int x = 0; x < ne; x++) {
			const int* E = W.me + 8 * (size_t)el[x];
			uint64_t m0 = 0, m1 = 0;
			for (int h = 0; h < 2; h++)
				if (mv[E[h]] == 0) {
					m0 |= W.col_used[2 * (size_t)E[h]];
					m1 |= W.col_used[2 * (size_t)E[h] + 1];
				}
			int c = ~m0 ? __builtin_ctzll(~m0) : (~m1 ? 64 + __builtin_ctzll(~m1) : CP_MAXC);
			col[x] = c;
			if (c == CP_MAXC) {
				any_serial = true;
				continue;
			}
			ncol = max(ncol, c + 1);
			for (int h = 0; h < 2; h++)
				if (mv[E[h]] == 0) W.col_used[2 * (size_t)E[h] + (c >> 6)] |= 1ull << (c & 63);
		}
		for (int x = 0; x < ne; x++) {
			const int* E = W.me + 8 * (size_t)el[x];
			for (int h = 0; h < 2; h++)
				if (mv[E[h]] == 0) {
					W.col_used[2 * (size_t)E[h]] = 0;
					W.col_used[2 * (size_t)E[h] + 1] = 0;
				}
		}
		// colour-major order (stable): counting sort into eorder
		const int nct = ncol + (any_serial ? 1 : 0);
		const int c0 = ebase + t;
		int* co = W.coff + c0;
		for (int c = 0; c <= nct; c++) co[c] = 0;
		for (int x = 0; x < ne; x++) co[(col[x] == CP_MAXC ? ncol : col[x]) + 1]++;
		for (int c = 0; c < nct; c++) co[c + 1] += co[c];
		int* eo = W.eorder + ebase;
		for (int x = 0; x < ne; x++) eo[co[col[x] == CP_MAXC ? ncol : col[x]]++] = x;
		for (int c = nct; c > 0; c--) co[c] = co[c - 1] + ebase;
		co[0] = ebase;
		W.icoff[t] = c0;
		W.ncol[t] = nct;
		W.serial[t] = any_serial ? 1 : 0;
		W.eoff[ebase] = T.off[t];
		for (int y = 0; y < ne; y++) {
			emit(el[eo[y]]);
			W.eoff[ebase + y + 1] = T.off[t] + nr;
		}
		// island-local indices of the dynamic bodies in order of first
		// appearance; non-dynamic ends get the zero column nbi
		int* bod = W.bod + T.body_off[t];
		int nbi = 0;
		for (int r = 0; r < nr; r++) {
			int k = rows[2 * r], kind = rows[2 * r + 1];
			if (kind == 1) continue;  // the normal row covered the contact
			int ends[2];
			int* slot;
			if (kind == 2) {
				ends[0] = W.ja[k]; ends[1] = W.jb[k];
				slot = W.jloc + 2 * (size_t)k;
			} else {
				ends[0] = (int)(W.ckey[k] >> 32); ends[1] = (int)(W.ckey[k] & 0xFFFFFFFFu);
				slot = W.cloc + 2 * (size_t)k;
			}
			for (int h = 0; h < 2; h++) {
				int i = ends[h];
				if (mv[i] != 0) {
					slot[h] = -1;
					continue;
				}
				if (W.loc[i] < 0) {
					W.loc[i] = nbi;
					bod[nbi++] = i;
				}
				slot[h] = W.loc[i];
			}
		}
		for (int r = 0; r < nr; r++) {
			int k = rows[2 * r], kind = rows[2 * r + 1];
			if (kind == 1) continue;
			int* slot = kind == 2 ? W.jloc + 2 * (size_t)k : W.cloc + 2 * (size_t)k;
			for (int h = 0; h < 2; h++)
				if (slot[h] < 0) slot[h] = nbi;
		}
		for (int x = 0; x < nbi; x++) W.loc[bod[x]] = -1;
		W.nbod[t] = nbi;
	}
	// holonomic systems in order of first appearance; linearization-error
	// columns: dynamic bodies their own index, non-dynamic ones a column of
	// this island (nb + 2 * joint_off + slot)
	const int jb = T.joint_off[t];
	int nsys = 0, nslot = 0;
	int* sys = W.sys + jb;
	int* slots = W.slots + 2 * (size_t)jb;
	for (int r = 0; r < nr; r++) {
		if (rows[2 * r + 1] != 2) continue;
		int j = rows[2 * r];
		int q = W.hol ? W.jsys[j] : -1;
		if (q >= 0 && W.sys_seen[q] != 1) {
			W.sys_seen[q] = 1;
			sys[nsys++] = q;
		}
		int ends[2] = {W.ja[j], W.jb[j]};
		for (int h = 0; h < 2; h++) {
			int i = ends[h], cidx = i;
			if (mv[i] != 0) {
				int s = 0;
				while (s < nslot && slots[s] != i) s++;
				if (s == nslot) slots[nslot++] = i;
				cidx = W.nb + 2 * jb + s;
			}
			W.jlin[2 * (size_t)j + h] = cidx;
		}
	}
	for (int s = 0; s < nsys; s++) W.sys_seen[sys[s]] = 0;
	W.sys_cnt[t] = nsys;
}

static inline int nblk256(int n) { return (n + 255) / 256; }

void launch_ib_keys(const uint64_t* ekey, int ne, const uint64_t* jgkey, int ng, uint64_t* keys, int* vals, hipStream_t s) {
	if (ne + ng) hipLaunchKernelGGL(k_ib_keys, dim3(nblk256(ne + ng)), dim3(256), 0, s, ekey, ne, jgkey, ng, keys, vals);
}
void launch_ib_heads(const uint64_t* skeys, int n, int* head, hipStream_t s) {
	hipLaunchKernelGGL(k_ib_heads, dim3(nblk256(n + 1)), dim3(256), 0, s, skeys, n, head);
}
void launch_ib_edges(const uint64_t* skeys, const int* svals, int n, const int* head, const int* midx, int ne, const int* efirst,
                     const int* ecnt, const int* live, const int* jgoff, const int* jord, const JointAct& jact, int* me,
                     uint64_t* adjk, int* adjv, hipStream_t s) {
	if (n)
		hipLaunchKernelGGL(k_ib_edges, dim3(nblk256(n)), dim3(256), 0, s, skeys, svals, n, head, midx, ne, efirst, ecnt, live, jgoff, jord,
		                   jact, me, adjk, adjv);
}
void launch_ib_aoff(const uint64_t* adjk, int n2, int nb, int* aoff, hipStream_t s) {
	hipLaunchKernelGGL(k_ib_aoff, dim3(nblk256(n2 + 1)), dim3(256), 0, s, adjk, n2, nb, aoff);
}
void launch_ib_components(const int* me, const int* midx, int n, int nb, const int* movement, const int* rank, const int* aoff, int* parent,
                          int* comp, IslCnt* acc, const SleepDev& S, hipStream_t s) {
	hipLaunchKernelGGL(k_ib_init, dim3(nblk256(nb)), dim3(256), 0, s, parent, acc, nb);
	if (n) hipLaunchKernelGGL(k_ib_union, dim3(nblk256(n)), dim3(256), 0, s, me, midx, n, movement, rank, parent);
	hipLaunchKernelGGL(k_ib_bodies, dim3(nblk256(nb)), dim3(256), 0, s, nb, movement, aoff, parent, comp, acc, S);
	if (n) hipLaunchKernelGGL(k_ib_edge_counts, dim3(nblk256(n)), dim3(256), 0, s, me, midx, n, movement, comp, S.enabled ? S.asleep : nullptr, acc);
}
void launch_ib_flags(int nb, const int* rord, const int* comp, const IslCnt* acc, int sleeping, IslCnt* in, hipStream_t s) {
	hipLaunchKernelGGL(k_ib_flags, dim3(nblk256(nb + 1)), dim3(256), 0, s, nb, rord, comp, acc, sleeping, in);
}
void launch_ib_tables(int nb, const int* rord, const IslCnt* in, const IslCnt* out, const IslTables& T, int* info, const int* comp,
                      const IslCnt* acc, const SleepDev& S, double* vel, double* angv, hipStream_t s) {
	hipLaunchKernelGGL(k_ib_tables, dim3(nblk256(nb + 1)), dim3(256), 0, s, nb, rord, in, out, T, info);
	if (S.sleeping) hipLaunchKernelGGL(k_ib_sleep, dim3(nblk256(nb)), dim3(256), 0, s, nb, comp, acc, S, vel, angv);
}
void launch_ib_walk(const IslWalk& W, int nisl, bool colored, hipStream_t s) {
	if (!nisl) return;
	if (colored) hipLaunchKernelGGL(k_ib_walk<true>, dim3(nblk256(nisl)), dim3(256), 0, s, W, nisl);
	else hipLaunchKernelGGL(k_ib_walk<false>, dim3(nblk256(nisl)), dim3(256), 0, s, W, nisl);
}

}  // namespace cpx
