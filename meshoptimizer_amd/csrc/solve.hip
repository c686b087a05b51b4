// The vertex solve of meshopt_simplifyWithUpdate (src/simplifier.cpp:2575-2599) on the device.
//
// solvePositions (src/simplifier.cpp:1675-1763) walks the primaries in index order, and each
// vertex reads the current positions of its welded neighbours: a lower-index neighbour has
// already been moved (or kept), a higher-index one has not.  The walk is replayed exactly in
// rounds, Kahn-style: every listed primary counts its listed lower-index neighbour entries
// (`waiting`); those with none form the first round's queue; a primary solved in round r takes
// one off the count of each listed higher-index neighbour entry, and the neighbour whose count
// reaches zero joins round r + 1's queue.  Two adjacent listed primaries are never in one round
// (the higher waits for the lower), so nothing a round reads is written in that round and each
// vertex sees precisely the positions the sequential walk would show it.  Rounds = the longest
// index-increasing path through listed neighbours (tens for subdivided meshes, hundreds for
// row-major grids); work = O(V + E) over all rounds.
//
// Most outcomes are decided before the rounds.  A primary's solved position and its error test
// depend on the vertex alone; only the two neighbourhood tests (radius and triangle flips) read
// the neighbours, and both only reject.  A lower-index neighbour whose own outcome is still open
// may hold its current or its solved position when the walk reaches the vertex, so the tests are
// evaluated over both: an outcome that is the same in every combination is the walk's, and the
// vertex is decided (keep or move) without waiting.  Two such passes (the second sees the first's
// decisions) leave only the primaries whose outcome hinges on an open neighbour, and the rounds
// replay the walk for them alone, reading a decided lower-index neighbour's final position.  A
// decided move is written to the positions only after the rounds, since every undecided primary
// below it must still see its current position.
//
// solveAttributes (1765-1809) and finalizeVertices (623-663) are independent per vertex.
#include <algorithm>

#include "kernels.h"
#include "prims.h"
#include "quadric.h"

namespace mhip
{

constexpr uint32_t kPending = 0xffffffffu;
static constexpr int kThreads = 256;

// a listed primary's outcome as far as decided before the rounds
enum : uint8_t
{
	kSpecOpen = 0, // hinges on a lower-index neighbour still open: solved in the rounds
	kSpecKeep = 1, // keeps its position (not listed, no solution, or rejected in every case)
	kSpecMove = 2, // moves to its solved position in every case
};

static void launch_solve_decide(hipStream_t s, Mesh& m, const SolveBuffers& b);

static inline unsigned int grid_for(size_t n)
{
	return div_up(n ? n : 1, kThreads);
}

__device__ __forceinline__ bool solve_pinned(uint8_t k)
{
	return k == kLocked || k == kSeam || k == kBorder;
}

// quadricSolve, src/simplifier.cpp:967-1022: LDL^T factorisation of the 3x3 block, the volume
// gradient appended as a fourth row (a linear constraint with a Lagrange multiplier, dropped
// when its pivot is tiny); false when one of the first three pivots is tiny
__device__ __forceinline__ bool quad_solve(V3& p, const Quad& Q, const Grad& GV)
{
	float eps = 1e-6f * Q.w;

	float d0 = Q.a00;
	float l10 = Q.a10 / d0;
	float l20 = Q.a20 / d0;
	float d1 = Q.a11 - Q.a10 * l10;
	float dl21 = Q.a21 - Q.a20 * l10;
	float l21 = dl21 / d1;
	float d2 = Q.a22 - Q.a20 * l20 - dl21 * l21;

	// forward substitution of -b, then the diagonal
	float y0 = -Q.b0;
	float y1 = -Q.b1 - l10 * y0;
	float y2 = -Q.b2 - l20 * y0 - l21 * y1;
	float z0 = y0 / d0;
	float z1 = y1 / d1;
	float z2 = y2 / d2;

	// constraint row
	float l30 = GV.gx / d0;
	float dl31 = GV.gy - GV.gx * l10;
	float l31 = dl31 / d1;
	float dl32 = GV.gz - GV.gx * l20 - dl31 * l21;
	float l32 = dl32 / d2;
	float d3 = 0.f - GV.gx * l30 - dl31 * l31 - dl32 * l32;
	float y3 = -GV.gw - l30 * y0 - l31 * y1 - l32 * y2;
	float lambda = fabsf(d3) > eps ? y3 / d3 : 0.f;

	// back substitution
	p.z = z2 - l32 * lambda;
	p.y = z1 - l21 * p.z - l31 * lambda;
	p.x = z0 - l10 * p.y - l20 * p.z - l30 * lambda;

	return fabsf(d0) > eps && fabsf(d1) > eps && fabsf(d2) > eps;
}

// quadricReduceAttributes, src/simplifier.cpp:1024-1055: one wedge's attribute quadric (scaled
// by the vertex weight) and its gradients (the B B^T / a Schur term) folded into Q
__device__ __forceinline__ void quad_reduce_attributes(Quad& Q, const Quad& A, const Grad* G, int n)
{
	float w = Q.w;
	Q.a00 += A.a00 * w;
	Q.a11 += A.a11 * w;
	Q.a22 += A.a22 * w;
	Q.a10 += A.a10 * w;
	Q.a20 += A.a20 * w;
	Q.a21 += A.a21 * w;
	Q.b0 += A.b0 * w;
	Q.b1 += A.b1 * w;
	Q.b2 += A.b2 * w;

	float iaw = A.w == 0 ? 0.f : w / A.w;

	for (int k = 0; k < n; ++k)
	{
		Grad g = G[k];
		Q.a00 -= (g.gx * g.gx) * iaw;
		Q.a11 -= (g.gy * g.gy) * iaw;
		Q.a22 -= (g.gz * g.gz) * iaw;
		Q.a10 -= (g.gx * g.gy) * iaw;
		Q.a20 -= (g.gx * g.gz) * iaw;
		Q.a21 -= (g.gy * g.gz) * iaw;
		Q.b0 -= (g.gx * g.gw) * iaw;
		Q.b1 -= (g.gy * g.gw) * iaw;
		Q.b2 -= (g.gz * g.gw) * iaw;
	}
}

// ------------------------------------------------------------------------------------------
// set-up: vertex_update (src/simplifier.cpp:2581-2588) and the list of primaries to solve

__global__ void k_solve_mark(const uint32_t* ib, size_t n, const uint32_t* remap, uint8_t* update)
{
	size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (j >= n)
		return;
	uint32_t v = ib[j];
	update[v] = 1;
	update[remap[v]] = 1;
}

__global__ void k_solve_flags(const uint8_t* update, const uint32_t* remap, const uint8_t* kind, size_t n, uint32_t* listed)
{
	size_t ii = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (ii >= n)
		return;
	uint32_t i = (uint32_t)ii;
	listed[i] = update[i] && remap[i] == i && !solve_pinned(kind[i]);
}

// wave-aggregated append of `item` (for the lanes with `take`) to a queue
__device__ __forceinline__ void queue_append(bool take, uint32_t item, uint32_t* count, uint32_t* queue)
{
	const uint64_t mask = __ballot(take);
	if (!mask)
		return;
	const uint32_t lane = __lane_id();
	const uint32_t leader = (uint32_t)__builtin_ctzll(mask);
	uint32_t base = 0;
	if (lane == leader)
		base = atomicAdd(count, (uint32_t)__popcll(mask));
	base = __shfl(base, (int)leader, 64);
	if (take)
		queue[base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull))] = item;
}

// open lower-index neighbour entries of every open primary (the decided ones need no round);
// the primaries with none are the first round's queue; sc->work[4] = open primaries.  Both
// counters take one atomic per workgroup (same-address atomics serialise in L2).
__global__ void __launch_bounds__(kThreads) k_solve_waiting(const uint8_t* state, const uint32_t* adj_off, const uint2* adj, size_t n, uint32_t* waiting,
                                                             uint32_t* queue0, uint32_t* counts, CallScalars* sc)
{
	__shared__ uint32_t s_take[kThreads / 64], s_listed[kThreads / 64], s_base;
	size_t ii = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	uint32_t i = (uint32_t)ii;
	bool mine = ii < n && state[i] == kSpecOpen;
	uint32_t w = 0;
	if (mine)
	{
		for (uint32_t e = adj_off[i], end = adj_off[i + 1]; e < end; ++e)
		{
			uint2 ed = adj[e];
			w += (ed.x < i && state[ed.x] == kSpecOpen) + (ed.y < i && state[ed.y] == kSpecOpen);
		}
	}
	// decided and unlisted vertices get a count no release can bring to zero (at most 2 per
	// corner entry)
	if (ii < n)
		waiting[i] = mine ? w : 0x80000000u;

	const bool take = mine && w == 0;
	const uint64_t tm = __ballot(take), lm = __ballot(mine);
	const uint32_t lane = __lane_id(), wave = threadIdx.x / 64;
	if (lane == 0)
	{
		s_take[wave] = (uint32_t)__popcll(tm);
		s_listed[wave] = (uint32_t)__popcll(lm);
	}
	__syncthreads();
	if (threadIdx.x == 0)
	{
		uint32_t tt = 0, lt = 0;
		for (int q = 0; q < kThreads / 64; ++q)
		{
			tt += s_take[q];
			lt += s_listed[q];
		}
		s_base = tt ? atomicAdd(&counts[0], tt) : 0u;
		if (lt)
			atomicAdd(&sc->work[4], lt);
	}
	__syncthreads();
	if (take)
	{
		uint32_t at = s_base + (uint32_t)__popcll(tm & ((1ull << lane) - 1ull));
		for (uint32_t q = 0; q < wave; ++q)
			at += s_take[q];
		queue0[at] = i;
	}
}

__global__ void k_solve_reset(CallScalars* sc)
{
	sc->work[4] = 0;
	sc->work[5] = sc->work[6] = sc->work[7] = 0;
}

void launch_solve_setup(hipStream_t s, Mesh& m, const SolveBuffers& b, size_t result_count)
{
	size_t n = m.vertex_count;
	k_solve_reset<<<1, 1, 0, s>>>(m.sc);
	if (!n)
		return;
	fill_u32(s, reinterpret_cast<uint32_t*>(b.update), 0, (n + 4) / 4);
	if (result_count)
		k_solve_mark<<<grid_for(result_count), kThreads, 0, s>>>(m.ib[m.ib_cur], result_count, m.remap, b.update);
	// updateEdgeAdjacency over the result (src/simplifier.cpp:2590-2591)
	launch_adjacency(s, m, result_count / 3);
	k_solve_flags<<<grid_for(n), kThreads, 0, s>>>(b.update, m.remap, m.kind, n, b.listed);
	launch_solve_decide(s, m, b);
	k_solve_waiting<<<grid_for(n), kThreads, 0, s>>>(b.state, m.adj_off, m.adj, n, b.waiting, b.queue[0], &m.sc->work[5], m.sc);
}

// ------------------------------------------------------------------------------------------
// one round of solvePositions (src/simplifier.cpp:1675-1763)

struct SolveRound
{
	uint32_t* waiting;
	uint32_t* queue[2];
	uint32_t* counts;
	float4* pos;
	const Quad* vq;
	const Grad* vg;
	const Quad* aq;
	const Grad* ag;
	int attr_count;
	const uint32_t* wedge;
	const uint32_t* adj_off;
	const uint2* adj;
	float4* cand;
	uint8_t* state;
	const uint32_t* listed;
};

// the solved position of listed primary i and the error test (src/simplifier.cpp:1697-1729 and
// 1750-1755): false when there is no solution or it raises the positional error too much
__device__ __forceinline__ bool solve_candidate(const SolveRound& a, uint32_t i, const V3& cur, V3& p)
{
	// solvePositions' system (src/simplifier.cpp:1705-1729): a point quadric regularises the
	// solution; with attributes the optimum minimises every wedge's attribute quadric at once,
	// under the volume gradient as a constraint
	const Quad Qv = load_quad(a.vq, i);
	Quad Q = Qv;
	Quad R;
	quad_from_point(R, cur.x, cur.y, cur.z, Qv.w * 1e-4f);
	quad_add(Q, R);
	Grad GV = {0.f, 0.f, 0.f, 0.f};
	if (a.attr_count)
	{
		uint32_t v = i;
		do
		{
			quad_reduce_attributes(Q, load_quad(a.aq, v), a.ag + (size_t)v * a.attr_count, a.attr_count);
			v = a.wedge[v];
		} while (v != i);
		if (a.vg)
			GV = a.vg[i];
	}
	if (!quad_solve(p, Q, GV))
		return false;
	return !(quad_error(Qv, p) > quad_error(Qv, cur) * 1.5f + 1e-6f);
}

// every vertex: its solved position when listed, kSpecOpen when that is a candidate move,
// kSpecKeep otherwise (unlisted vertices never move)
__global__ void k_solve_candidates(SolveRound a, size_t n)
{
	size_t ii = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (ii >= n)
		return;
	const uint32_t i = (uint32_t)ii;
	uint8_t st = kSpecKeep;
	if (a.listed[i])
	{
		const V3 cur = load_pos(a.pos, i);
		V3 p = cur;
		if (solve_candidate(a, i, cur, p))
			st = kSpecOpen;
		a.cand[i] = make_float4(p.x, p.y, p.z, 0.f);
	}
	a.state[i] = st;
}

// the positions corner u may hold when the walk reaches primary i: its current one, or a
// lower-index primary's decided one, or both for a lower-index primary still open
__device__ __forceinline__ int corner_options(const SolveRound& a, uint32_t u, uint32_t i, V3 opt[2])
{
	const uint8_t st = u < i ? a.state[u] : kSpecKeep;
	if (st == kSpecMove)
	{
		opt[0] = load_pos(a.cand, u);
		return 1;
	}
	opt[0] = load_pos(a.pos, u);
	if (st == kSpecOpen)
	{
		opt[1] = load_pos(a.cand, u);
		return 2;
	}
	return 1;
}

// the position the walk shows at primary i for corner u once every open lower-index primary is
// done: a lower-index primary decided to move at its solved position, any other at its current
__device__ __forceinline__ V3 walk_pos(const SolveRound& a, uint32_t u, uint32_t i)
{
	return load_pos(u < i && a.state[u] == kSpecMove ? a.cand : a.pos, u);
}

// getNeighborhoodRadius's term: a NaN distance never enters its running maximum, which starts
// at 0, so it counts as 0
__device__ __forceinline__ float radius_term(const V3& q, const V3& cur)
{
	const float d = (q.x - cur.x) * (q.x - cur.x) + (q.y - cur.y) * (q.y - cur.y) + (q.z - cur.z) * (q.z - cur.z);
	return d == d ? d : 0.f;
}

// one pass over the open primaries (one thread each): the radius test over the smallest and the
// largest radius the combinations allow, and the flip test of every corner entry over its
// combinations (src/simplifier.cpp:1731-1748).  A neighbour decided in this same pass may be read
// as open or as decided: either way its true position is among the ones examined.
__global__ void k_solve_speculate(SolveRound a, size_t n)
{
	size_t ii = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (ii >= n)
		return;
	const uint32_t i = (uint32_t)ii;
	if (a.state[i] != kSpecOpen)
		return;
	const V3 cur = load_pos(a.pos, i), p = load_pos(a.cand, i);
	const float dp = (p.x - cur.x) * (p.x - cur.x) + (p.y - cur.y) * (p.y - cur.y) + (p.z - cur.z) * (p.z - cur.z);
	float lo2 = 0.f, hi2 = 0.f;
	bool flips_always = false, flips_maybe = false;
	for (uint32_t e = a.adj_off[i], end = a.adj_off[i + 1]; e < end; ++e)
	{
		const uint2 ed = a.adj[e];
		V3 qa[2], qb[2];
		const int na = corner_options(a, ed.x, i, qa), nb = corner_options(a, ed.y, i, qb);
		float ra_lo = radius_term(qa[0], cur), ra_hi = ra_lo, rb_lo = radius_term(qb[0], cur), rb_hi = rb_lo;
		if (na == 2)
		{
			const float r = radius_term(qa[1], cur);
			ra_lo = r < ra_lo ? r : ra_lo;
			ra_hi = r > ra_hi ? r : ra_hi;
		}
		if (nb == 2)
		{
			const float r = radius_term(qb[1], cur);
			rb_lo = r < rb_lo ? r : rb_lo;
			rb_hi = r > rb_hi ? r : rb_hi;
		}
		lo2 = lo2 < ra_lo ? ra_lo : lo2;
		lo2 = lo2 < rb_lo ? rb_lo : lo2;
		hi2 = hi2 < ra_hi ? ra_hi : hi2;
		hi2 = hi2 < rb_hi ? rb_hi : hi2;
		bool all = true, any = false;
		for (int x = 0; x < na; ++x)
			for (int y = 0; y < nb; ++y)
			{
				const bool f = triangle_flips(qa[x], qb[y], cur, p);
				all &= f;
				any |= f;
			}
		flips_always |= all;
		flips_maybe |= any;
	}
	// the radius test passes when !(dp > nr * nr), nr = sqrtf(radius), monotone in the radius
	const float nlo = sqrtf(lo2), nhi = sqrtf(hi2);
	if (dp > nhi * nhi || flips_always)
		a.state[i] = kSpecKeep;
	else if (!(dp > nlo * nlo) && !flips_maybe)
		a.state[i] = kSpecMove;
}

// the decided moves, once the rounds are done
__global__ void k_solve_apply(const uint8_t* state, const float4* cand, size_t n, float4* pos)
{
	size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n && state[i] == kSpecMove)
		pos[i] = cand[i];
}

static SolveRound solve_round_args(Mesh& m, const SolveBuffers& b)
{
	SolveRound a;
	a.waiting = b.waiting;
	a.queue[0] = b.queue[0];
	a.queue[1] = b.queue[1];
	a.counts = &m.sc->work[5];
	a.pos = m.pos;
	a.vq = m.vq;
	a.vg = m.vg;
	a.aq = m.aq;
	a.ag = m.ag;
	a.attr_count = (int)m.attr_count;
	a.wedge = m.wedge;
	a.adj_off = m.adj_off;
	a.adj = m.adj;
	a.cand = b.cand;
	a.state = b.state;
	a.listed = b.listed;
	return a;
}

static constexpr int kSpeculationPasses = 2;

static void launch_solve_decide(hipStream_t s, Mesh& m, const SolveBuffers& b)
{
	const size_t n = m.vertex_count;
	const SolveRound a = solve_round_args(m, b);
	k_solve_candidates<<<grid_for(n), kThreads, 0, s>>>(a, n);
	for (int k = 0; k < kSpeculationPasses; ++k)
		k_solve_speculate<<<grid_for(n), kThreads, 0, s>>>(a, n);
}

// neighbourhood checks of a solved position over `n` corner entries (src/simplifier.cpp:1734-1748):
// getNeighborhoodRadius (1252-1276) and hasTriangleFlips for a point move (1234-1250)
__device__ __forceinline__ void sweep_entry(const V3& q0, const V3& q1, const V3& cur, const V3& p, float& far2, bool& flips)
{
	float d0 = (q0.x - cur.x) * (q0.x - cur.x) + (q0.y - cur.y) * (q0.y - cur.y) + (q0.z - cur.z) * (q0.z - cur.z);
	float d1 = (q1.x - cur.x) * (q1.x - cur.x) + (q1.y - cur.y) * (q1.y - cur.y) + (q1.z - cur.z) * (q1.z - cur.z);
	far2 = far2 < d0 ? d0 : far2;
	far2 = far2 < d1 ? d1 : far2;
	flips |= triangle_flips(q0, q1, cur, p);
}

// round r: decide the queued (open) primaries, one wavefront per primary, and release their
// higher-index open neighbours into round r + 1's queue.  Queue lengths rotate over three words:
// round r reads counts[r % 3], appends to counts[(r + 1) % 3] and clears counts[(r + 2) % 3]
// (round r - 1's, read by nobody after it).
//
// A queued primary's solved position passed the error test (k_solve_candidates); the lanes share
// its corner entries.  A corner is read where the walk would show it: a lower-index primary
// decided to move at its solved position (written to the positions only after the rounds), every
// other one at the current position (an open lower-index primary was moved in an earlier round).
// The release does not depend on the outcome (the neighbour reads the new position only in a
// later launch), so each lane issues the atomics of its first entry before the tests and uses
// their results after them.  Both neighbourhood checks only reject, so evaluating both changes
// no outcome; the radius is a maximum of non-negative values that never takes a NaN (a maximum
// of bit patterns, order-free) and the flip test an any().
__global__ void __launch_bounds__(kThreads) k_solve_round(SolveRound a, uint32_t round)
{
	const uint32_t next = (round + 1) % 3;
	if (blockIdx.x == 0 && threadIdx.x == 0)
		a.counts[(round + 2) % 3] = 0;
	const uint32_t count = a.counts[round % 3];
	const uint32_t* queue = a.queue[round & 1];
	uint32_t* out = a.queue[(round + 1) & 1];
	const uint32_t lane = __lane_id();
	const uint32_t waves = gridDim.x * (blockDim.x / 64);

	for (uint32_t k = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64; k < count; k += waves)
	{
		const uint32_t i = queue[k];
		const uint32_t begin = a.adj_off[i], end = a.adj_off[i + 1];
		const V3 cur = load_pos(a.pos, i), p = load_pos(a.cand, i);

		// this lane's first corner entry: release its higher-index ends now (decided and unlisted
		// vertices hold a count that never reaches zero), load the positions the sweep needs
		const uint32_t e0 = begin + lane;
		const bool has0 = e0 < end;
		const uint2 ed0 = has0 ? a.adj[e0] : make_uint2(i, i);
		const uint32_t old0 = ed0.x > i ? atomicSub(&a.waiting[ed0.x], 1u) : 0u;
		const uint32_t old1 = ed0.y > i ? atomicSub(&a.waiting[ed0.y], 1u) : 0u;
		const V3 q0 = walk_pos(a, ed0.x, i), q1 = walk_pos(a, ed0.y, i);

		bool moved = true;
		{
			float far2 = 0.f;
			bool flips = false;
			if (has0)
				sweep_entry(q0, q1, cur, p, far2, flips);
			for (uint32_t e = e0 + 64; e < end; e += 64)
			{
				uint2 ed = a.adj[e];
				sweep_entry(walk_pos(a, ed.x, i), walk_pos(a, ed.y, i), cur, p, far2, flips);
			}
			uint32_t far_bits = __float_as_uint(far2);
			for (int o = 32; o > 0; o >>= 1)
			{
				uint32_t t = __shfl_xor(far_bits, o, 64);
				far_bits = t > far_bits ? t : far_bits;
			}
			float nr = sqrtf(__uint_as_float(far_bits));
			float dp = (p.x - cur.x) * (p.x - cur.x) + (p.y - cur.y) * (p.y - cur.y) + (p.z - cur.z) * (p.z - cur.z);
			moved = !(dp > nr * nr) && !__ballot(flips);
		}
		if (moved && lane == 0)
			a.pos[i] = make_float4(p.x, p.y, p.z, 0.f);

		queue_append(old0 == 1u, ed0.x, &a.counts[next], out);
		queue_append(old1 == 1u, ed0.y, &a.counts[next], out);

		// corner entries past the first 64 (high-valence vertices)
		for (uint32_t base = begin + 64; base < end; base += 64)
		{
			const uint32_t e = base + lane;
			uint32_t r0 = kNone, r1 = kNone;
			if (e < end)
			{
				uint2 ed = a.adj[e];
				if (ed.x > i && atomicSub(&a.waiting[ed.x], 1u) == 1u)
					r0 = ed.x;
				if (ed.y > i && atomicSub(&a.waiting[ed.y], 1u) == 1u)
					r1 = ed.y;
			}
			queue_append(r0 != kNone, r0, &a.counts[next], out);
			queue_append(r1 != kNone, r1, &a.counts[next], out);
		}
	}
}

void launch_solve_rounds(hipStream_t s, Mesh& m, const SolveBuffers& b, uint32_t listed, uint32_t first, uint32_t rounds)
{
	if (!listed)
		return;
	const SolveRound a = solve_round_args(m, b);
	// one wavefront per queued primary, striding over the queue
	const unsigned int grid = (unsigned int)std::min<size_t>((listed + 3) / 4, 1024);
	for (uint32_t r = 0; r < rounds; ++r)
		k_solve_round<<<grid, kThreads, 0, s>>>(a, first + r);
}

// ------------------------------------------------------------------------------------------
// after the rounds

// wedges take their primary's final position (src/simplifier.cpp:1691-1695)
__global__ void k_solve_wedges(const uint8_t* update, const uint32_t* remap, const uint8_t* kind, size_t n, float4* pos)
{
	size_t ii = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (ii >= n)
		return;
	uint32_t i = (uint32_t)ii;
	if (update[i] && !solve_pinned(kind[i]) && remap[i] != i)
		pos[i] = pos[remap[i]];
}

// solveAttributes, src/simplifier.cpp:1765-1809: every wedge's attributes from its gradients at
// the solved position; a complex vertex whose wedges shared a value keeps it shared, taking
// the heaviest wedge's
__global__ void k_solve_attributes(const uint8_t* update, const uint32_t* remap, const uint32_t* wedge, const uint8_t* kind, const float4* pos,
                                   const Quad* aq, const Grad* ag, int n_attr, size_t n, float* va)
{
	size_t ii = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (ii >= n)
		return;
	uint32_t i = (uint32_t)ii;
	if (!update[i] || remap[i] != i)
		return;
	const V3 p = load_pos(pos, i);
	const size_t A = (size_t)n_attr;

	for (size_t k = 0; k < A; ++k)
	{
		uint32_t shared = kNone;
		if (kind[i] == kComplex)
		{
			shared = i;
			for (uint32_t v = wedge[i]; v != i; v = wedge[v])
			{
				if (va[v * A + k] != va[i * A + k])
					shared = kNone;
				else if (shared != kNone && aq[v].w > aq[shared].w)
					shared = v;
			}
		}

		uint32_t v = i;
		do
		{
			uint32_t src = shared == kNone ? v : shared;
			float w = aq[src].w;
			Grad G = ag[src * A + k];
			float iw = w == 0 ? 0.f : 1.f / w;
			va[v * A + k] = (G.gx * p.x + G.gy * p.y + G.gz * p.z + G.gw) * iw;
			v = wedge[v];
		} while (v != i);
	}
}

// finalizeVertices, src/simplifier.cpp:623-663: back to the caller's space and unweighted
// attributes, skipping externally locked vertices (and the positions of locked kinds)
__global__ void k_solve_finalize(const uint8_t* update, const uint8_t* kind, const float4* pos, const float* va, int n_attr, size_t n,
                                 const CallScalars* sc, SolveOutput o)
{
	size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n || !update[i])
		return;
	size_t ri = o.sparse_remap ? o.sparse_remap[i] : i;
	if (o.vertex_lock && (o.vertex_lock[ri] & kVertexLock))
		return;

	if (kind[i] != kLocked)
	{
		float4 p = pos[i];
		float scale = sc->extent;
		float* out = o.positions + ri * o.position_stride;
		out[0] = p.x * scale + sc->lo[0];
		out[1] = p.y * scale + sc->lo[1];
		out[2] = p.z * scale + sc->lo[2];
	}

	for (int k = 0; k < n_attr; ++k)
	{
		uint32_t rk = o.map.index[k];
		o.attributes[ri * o.attribute_stride + rk] = va[i * (size_t)n_attr + k] / o.map.weight[k];
	}
}

void launch_solve_finish(hipStream_t s, Mesh& m, const SolveBuffers& b, const SolveOutput& o)
{
	size_t n = m.vertex_count;
	if (!n)
		return;
	k_solve_apply<<<grid_for(n), kThreads, 0, s>>>(b.state, b.cand, n, m.pos);
	k_solve_wedges<<<grid_for(n), kThreads, 0, s>>>(b.update, m.remap, m.kind, n, m.pos);
	if (m.attr_count)
		k_solve_attributes<<<grid_for(n), kThreads, 0, s>>>(b.update, m.remap, m.wedge, m.kind, m.pos, m.aq, m.ag, (int)m.attr_count, n, m.va);
	k_solve_finalize<<<grid_for(n), kThreads, 0, s>>>(b.update, m.kind, m.pos, m.va, (int)m.attr_count, n, m.sc, o);
}

} // namespace mhip
