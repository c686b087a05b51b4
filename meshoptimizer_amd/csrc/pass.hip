// Per-pass kernels of the edge-collapse simplifier (src/simplifier.cpp:2493-2541):
// welded adjacency, pick, rank, sort, the priority-round greedy commit with its prefix-scan
// stop, quadric update, edge-loop remap and index compaction.
#include "kernels.h"
#include "prims.h"
#include "quadric.h"

namespace mhip
{

static constexpr int kThreads = 256;

static inline unsigned int grid_for(size_t n)
{
	return div_up(n ? n : 1, kThreads);
}

// kCanCollapse / kHasOpposite, src/simplifier.cpp:312-331
__constant__ uint8_t c_can_collapse[5][5] = {
    {1, 1, 1, 1, 1},
    {0, 1, 0, 0, 1},
    {0, 0, 1, 0, 1},
    {0, 0, 0, 1, 1},
    {0, 0, 0, 0, 0},
};

__constant__ uint8_t c_has_opposite[5][5] = {
    {1, 1, 1, 1, 1},
    {1, 0, 1, 0, 0},
    {1, 1, 1, 0, 1},
    {1, 0, 0, 0, 0},
    {1, 0, 1, 0, 0},
};

// ------------------------------------------------------------------------------------------
// welded corner CSR for this pass (src/simplifier.cpp:52-109).  Segment order is free: every
// consumer (open-edge tests, flip tests, neighbourhood scans) is a count or an any().  One atomic
// pass counts the corners per class and hands each corner its slot in the class's segment; after
// the scan the fill places the corners without atomics.

// the live triangle count: the host's when `live` is null, else the device word (at most the
// host's, which sizes the grid)
__device__ __forceinline__ size_t live_tris(size_t tris_max, const uint32_t* live)
{
	if (!live)
		return tris_max;
	size_t n = *live;
	return n < tris_max ? n : tris_max;
}

__global__ void k_adj_count(const uint32_t* ib, size_t tris_max, const uint32_t* live, const uint32_t* remap, uint32_t* counts, uint32_t* slot)
{
	size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (j < 3 * live_tris(tris_max, live))
		slot[j] = atomicAdd(&counts[remap[ib[j]]], 1u);
}

__global__ void k_adj_fill(const uint32_t* ib, size_t tris_max, const uint32_t* live, const uint32_t* remap, const uint32_t* off, const uint32_t* slot,
                           uint2* adj)
{
	size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= live_tris(tris_max, live))
		return;
	uint32_t a = remap[ib[t * 3 + 0]], b = remap[ib[t * 3 + 1]], c = remap[ib[t * 3 + 2]];
	adj[off[a] + slot[t * 3 + 0]] = make_uint2(b, c);
	adj[off[b] + slot[t * 3 + 1]] = make_uint2(c, a);
	adj[off[c] + slot[t * 3 + 2]] = make_uint2(a, b);
}

// the corner slots go to m.keys, free whenever the adjacency is rebuilt (before pick, after the
// pass loop)
void launch_adjacency(hipStream_t s, Mesh& m, size_t tri_count, const uint32_t* live)
{
	size_t n = m.vertex_count;
	fill_u32(s, m.adj_off, 0, n + 1);
	if (tri_count)
		k_adj_count<<<grid_for(tri_count * 3), kThreads, 0, s>>>(m.ib[m.ib_cur], tri_count, live, m.remap, m.adj_off, m.keys);
	scan_exclusive(s, m.adj_off, m.adj_off, n, nullptr, m.adj_off + n, m.temp);
	if (tri_count)
		k_adj_fill<<<grid_for(tri_count), kThreads, 0, s>>>(m.ib[m.ib_cur], tri_count, live, m.remap, m.adj_off, m.keys, m.adj);
}

// ------------------------------------------------------------------------------------------
// pickEdgeCollapses, src/simplifier.cpp:1310-1372: candidates of one triangle in edge order

__device__ __forceinline__ int pick_triangle(const uint32_t* ib, size_t t, const uint32_t* remap, const uint8_t* kind, const uint32_t* loop,
                                             const uint32_t* loopback, uint32_t v0[3], uint32_t v1[3], uint32_t bidi[3])
{
	uint32_t tri[3] = {ib[t * 3 + 0], ib[t * 3 + 1], ib[t * 3 + 2]};
	int count = 0;
#pragma unroll
	for (int e = 0; e < 3; ++e)
	{
		uint32_t i0 = tri[e];
		uint32_t i1 = tri[e == 2 ? 0 : e + 1];
		uint32_t r0 = remap[i0], r1 = remap[i1];
		if (r0 == r1)
			continue;
		uint8_t k0 = kind[i0], k1 = kind[i1];
		uint8_t c01 = c_can_collapse[k0][k1], c10 = c_can_collapse[k1][k0];
		if (!(c01 | c10))
			continue;
		if (c_has_opposite[k0][k1] && r1 > r0)
			continue;
		if ((k0 == kBorder || k0 == kSeam) && k1 != kManifold && loop[i0] != i1)
			continue;
		if ((k1 == kBorder || k1 == kSeam) && k0 != kManifold && loopback[i1] != i0)
			continue;
		if (c01 & c10)
		{
			v0[count] = i0;
			v1[count] = i1;
			bidi[count] = 1;
		}
		else
		{
			v0[count] = c01 ? i0 : i1;
			v1[count] = c01 ? i1 : i0;
			bidi[count] = 0;
		}
		count++;
	}
	return count;
}

__global__ void k_pick_count(const uint32_t* ib, size_t tris_max, const uint32_t* live, const uint32_t* remap, const uint8_t* kind, const uint32_t* loop,
                             const uint32_t* loopback, uint32_t* counts)
{
	size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= live_tris(tris_max, live))
		return;
	uint32_t v0[3], v1[3], bd[3];
	counts[t] = (uint32_t)pick_triangle(ib, t, remap, kind, loop, loopback, v0, v1, bd);
}

// triangles are taken while `written + 3 <= capacity` (src/simplifier.cpp:1318-1320)
__global__ void k_pick_write(const uint32_t* ib, size_t tris_max, const uint32_t* live, const uint32_t* remap, const uint8_t* kind, const uint32_t* loop,
                             const uint32_t* loopback, const uint32_t* offsets, const uint32_t* capacity_word, uint32_t* c_v0, uint32_t* c_v1,
                             float* c_err, uint32_t* out_count)
{
	size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	const size_t tris = live_tris(tris_max, live);
	if (t >= tris)
		return;
	const uint32_t capacity = *capacity_word;
	uint32_t base = offsets[t];
	if ((uint64_t)base + 3 > capacity)
		return;
	uint32_t v0[3], v1[3], bd[3];
	int count = pick_triangle(ib, t, remap, kind, loop, loopback, v0, v1, bd);
	for (int k = 0; k < count; ++k)
	{
		c_v0[base + k] = v0[k];
		c_v1[base + k] = v1[k];
		c_err[base + k] = __uint_as_float(bd[k]); // the bidirectional flag until ranked
	}
	// offsets are non-decreasing, so the processed triangles form a prefix; its last one writes C
	if (t + 1 == tris || (uint64_t)offsets[t + 1] + 3 > capacity)
		*out_count = base + (uint32_t)count;
}

void launch_pick(hipStream_t s, Mesh& m, size_t tri_count, const uint32_t* live)
{
	uint32_t* loop = m.loop[m.loop_cur];
	uint32_t* loopback = m.loopback[m.loop_cur];
	HIP_CHECK(hipMemsetAsync(&m.sc->candidates, 0, sizeof(uint32_t), s));
	if (!tri_count)
		return;
	k_pick_count<<<grid_for(tri_count), kThreads, 0, s>>>(m.ib[m.ib_cur], tri_count, live, m.remap, m.kind, loop, loopback, m.tri_cnt);
	scan_exclusive(s, m.tri_cnt, m.tri_off, tri_count, live, nullptr, m.temp);
	k_pick_write<<<grid_for(tri_count), kThreads, 0, s>>>(m.ib[m.ib_cur], tri_count, live, m.remap, m.kind, loop, loopback, m.tri_off,
	                                                      &m.sc->pick_capacity, m.c_v0, m.c_v1, m.c_err, &m.sc->candidates);
}

// boundEdgeCollapses (src/simplifier.cpp:1292-1308) from the dual count on the device:
// index_count - dual / 2 + 3, so the first pick needs no read-back
__global__ void k_pick_capacity(CallScalars* sc, uint64_t index_count)
{
	const uint64_t c = index_count - sc->dual_count / 2 + 3;
	sc->pick_capacity = c < 0xffffffffull ? (uint32_t)c : 0xffffffffu;
}

void launch_pick_capacity(hipStream_t s, Mesh& m, size_t index_count)
{
	k_pick_capacity<<<1, 1, 0, s>>>(m.sc, (uint64_t)index_count);
}

// ------------------------------------------------------------------------------------------
// rankEdgeCollapses, src/simplifier.cpp:1374-1450

__device__ __forceinline__ uint32_t complex_target(uint32_t v, uint32_t target, const uint32_t* remap, const uint32_t* loop, const uint32_t* loopback)
{
	// getComplexTarget, src/simplifier.cpp:1278-1290
	uint32_t r = remap[target];
	uint32_t l = loop[v];
	if (l != kNone && remap[l] == r)
		return l;
	uint32_t lb = loopback[v];
	if (lb != kNone && remap[lb] == r)
		return lb;
	return target;
}

struct RankArgs
{
	const float4* pos;
	const float* va;
	const Quad* vq;
	const Quad* aq;
	const Grad* ag;
	int attr_count;
	const uint32_t* remap;
	const uint32_t* wedge;
	const uint8_t* kind;
	const uint32_t* loop;
	const uint32_t* loopback;
};

__device__ __forceinline__ float attr_error(const RankArgs& a, uint32_t v, uint32_t t)
{
	const int A = a.attr_count;
	return quad_error_attr(load_quad(a.aq, v), a.ag + (size_t)v * A, A, load_pos(a.pos, t), a.va + (size_t)t * A);
}

__global__ void k_rank(RankArgs a, size_t count, uint32_t* c_v0, uint32_t* c_v1, float* c_err, uint32_t* keys, uint32_t* vals)
{
	size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (c >= count)
		return;
	uint32_t i0 = c_v0[c], i1 = c_v1[c];
	bool bidi = __float_as_uint(c_err[c]) != 0;

	float ei = quad_error(load_quad(a.vq, a.remap[i0]), load_pos(a.pos, i1));
	float ej = bidi ? quad_error(load_quad(a.vq, a.remap[i1]), load_pos(a.pos, i0)) : FLT_MAX;

	if (a.attr_count)
	{
		ei += attr_error(a, i0, i1);
		ej += bidi ? attr_error(a, i1, i0) : 0;

		uint8_t k0 = a.kind[i0];
		if (k0 == kSeam)
		{
			uint32_t s0 = a.wedge[i0];
			uint32_t s1 = a.loop[i0] == i1 ? a.loopback[s0] : a.loop[s0];
			s1 = (s1 != kNone) ? s1 : a.wedge[i1];

			ei += attr_error(a, s0, s1);
			ej += bidi ? attr_error(a, s1, s0) : 0;
		}
		else
		{
			if (k0 == kComplex)
				for (uint32_t v = a.wedge[i0]; v != i0; v = a.wedge[v])
					ei += attr_error(a, v, complex_target(v, i1, a.remap, a.loop, a.loopback));

			if (a.kind[i1] == kComplex && bidi)
				for (uint32_t v = a.wedge[i1]; v != i1; v = a.wedge[v])
					ej += attr_error(a, v, complex_target(v, i0, a.remap, a.loop, a.loopback));
		}
	}

	bool rev = bidi & (ej < ei);
	float err = ej < ei ? ej : ei;
	c_v0[c] = rev ? i1 : i0;
	c_v1[c] = rev ? i0 : i1;
	c_err[c] = err;
	keys[c] = collapse_sort_key(err);
	vals[c] = (uint32_t)c;
}

__global__ void k_gather_ranked(const uint32_t* order, size_t count, const uint32_t* c_v0, const uint32_t* c_v1, const float* c_err,
                                const uint32_t* remap, uint32_t* r_v0, uint32_t* r_v1, uint32_t* r_r0, uint32_t* r_r1, float* r_err, uint8_t* status)
{
	size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= count)
		return;
	status[k] = kUndecided;
	uint32_t c = order[k];
	uint32_t v0 = c_v0[c], v1 = c_v1[c];
	r_v0[k] = v0;
	r_v1[k] = v1;
	r_r0[k] = remap[v0];
	r_r1[k] = remap[v1];
	r_err[k] = c_err[c];
}

void launch_rank_sort(hipStream_t s, Mesh& m, size_t cand_count)
{
	if (!cand_count)
		return;
	RankArgs a;
	a.pos = m.pos;
	a.va = m.va;
	a.vq = m.vq;
	a.aq = m.aq;
	a.ag = m.ag;
	a.attr_count = (int)m.attr_count;
	a.remap = m.remap;
	a.wedge = m.wedge;
	a.kind = m.kind;
	a.loop = m.loop[m.loop_cur];
	a.loopback = m.loopback[m.loop_cur];
	k_rank<<<grid_for(cand_count), kThreads, 0, s>>>(a, cand_count, m.c_v0, m.c_v1, m.c_err, m.keys, m.vals);
	// sortEdgeCollapses (src/simplifier.cpp:1452-1496): stable by the 12-bit key, ties in pick order
	radix_sort_pairs(s, m.keys, m.vals, m.keys_alt, m.vals_alt, cand_count, nullptr, 12, m.temp);
	k_gather_ranked<<<grid_for(cand_count), kThreads, 0, s>>>(m.vals, cand_count, m.c_v0, m.c_v1, m.c_err, m.remap, m.r_v0, m.r_v1, m.r_r0,
	                                                          m.r_r1, m.r_err, m.status);
}

// ------------------------------------------------------------------------------------------
// performEdgeCollapses (src/simplifier.cpp:1498-1636) as priority rounds.
//
// Rank k's decision depends only on lower ranks j that share an endpoint class (the lock
// test) or whose source class is a welded neighbour of k's source class (the flip test reads
// collapse_remap only there; only sources are written, and the neighbour relation is
// symmetric).  Round e: an undecided rank whose endpoints are locked is skipped (a lock can
// only come from a lower commit); one that was the minimum published at both endpoints in
// round e-1 and is below every source-role entry of e-1 in its neighbourhood decides now
// (flip-rejected or committed); the rest publish their rank for round e as 64-bit minima
// tagged with ~e in the high word, into the buffer of parity e, and stay on the work list.  A
// class keeps one minimum per role (source, target), so a rank publishes with two atomics, and
// the minimum over both roles is the smaller of the pair.
// Ranks that decide in one round never touch each other's state.

__device__ __forceinline__ uint64_t round_key(uint32_t epoch, uint32_t k)
{
	return ((uint64_t)(~epoch) << 32) | k;
}

// hasTriangleFlips, src/simplifier.cpp:1200-1232
__device__ __forceinline__ bool collapse_flips(const uint32_t* off, const uint2* adj, const float4* pos, const uint32_t* cremap, uint32_t i0, uint32_t i1)
{
	V3 v0 = load_pos(pos, i0), v1 = load_pos(pos, i1);
	for (uint32_t p = off[i0], e = off[i0 + 1]; p < e; ++p)
	{
		uint2 nb = adj[p];
		uint32_t a = cremap[nb.x];
		uint32_t b = cremap[nb.y];
		if (a == i1 || b == i1 || a == b)
			continue;
		if (triangle_flips(load_pos(pos, a), load_pos(pos, b), v0, v1))
			return true;
	}
	return false;
}

// commit writes of one collapse, src/simplifier.cpp:1573-1611
__device__ __forceinline__ void commit_collapse(uint32_t* cremap, uint32_t i0, uint32_t i1, const uint8_t* kind, const uint32_t* wedge,
                                                const uint32_t* remap, const uint32_t* loop, const uint32_t* loopback)
{
	uint8_t k = kind[i0];
	if (k == kComplex)
	{
		uint32_t v = i0;
		do
		{
			cremap[v] = complex_target(v, i1, remap, loop, loopback);
			v = wedge[v];
		} while (v != i0);
	}
	else if (k == kSeam)
	{
		uint32_t s0 = wedge[i0];
		uint32_t s1 = loop[i0] == i1 ? loopback[s0] : loop[s0];
		s1 = (s1 != kNone) ? s1 : wedge[i1];
		cremap[i0] = i1;
		cremap[s0] = s1;
	}
	else
		cremap[i0] = i1;
}

enum : uint8_t
{
	kFlipPass = 4, // flip test passes, now and for good; commits once lowest at its endpoints
	kFlipFail = 5, // flip test fails for good; resolved as skipped or rejected at the end
};

struct RoundState
{
	const uint32_t* r_v0;
	const uint32_t* r_v1;
	const uint32_t* adj_off;
	const uint2* adj;
	const float4* pos;
	const uint8_t* kind;
	const uint32_t* wedge;
	const uint32_t* remap;
	const uint32_t* loop;
	const uint32_t* loopback;
	const uint32_t* src_off;  // candidate ranks grouped by source class
	const uint32_t* src_rank;
	const uint32_t* src_val;  // what each such candidate's commit writes at the class primary
	uint32_t* cremap;
	uint32_t* lock_rank;      // rank of the commit that locked a class (kNone: unlocked)
	uint8_t* status;
	uint64_t* pend[2];        // per class, the minimum rank published in a round (tagged with ~round)
	                          // over ranks whose source is the class (entry 2c) and over ranks
	                          // whose target is the class (entry 2c + 1); the minimum over both
	                          // roles is the smaller of the two
	uint4* list[2];           // work-list entries: (rank, source class, target class, 0)
	CallScalars* sc;
	uint64_t seg_mul;         // list segment of class c: (c * seg_mul) >> 32, below kListSegments
};

// classes are numbered close to their neighbours in most meshes (grids, subdivided and
// locality-optimised meshes), so a segment's ranks touch mostly one range of the class arrays
__device__ __forceinline__ uint32_t list_segment(const RoundState& a, uint32_t c)
{
	return (uint32_t)(((uint64_t)c * a.seg_mul) >> 32);
}

// the value rank r's commit writes into collapse_remap[p], p the primary of its source class
// (src/simplifier.cpp:1573-1605)
__device__ __forceinline__ uint32_t commit_value_at(const RoundState& a, uint32_t r, uint32_t p)
{
	uint32_t i0 = a.r_v0[r], i1 = a.r_v1[r];
	uint8_t kd = a.kind[i0];
	if (kd == kComplex)
		return complex_target(p, i1, a.remap, a.loop, a.loopback);
	if (kd == kSeam)
	{
		uint32_t s0 = a.wedge[i0];
		uint32_t s1 = a.loop[i0] == i1 ? a.loopback[s0] : a.loop[s0];
		s1 = (s1 != kNone) ? s1 : a.wedge[i1];
		uint32_t v = p;
		if (i0 == p)
			v = i1;
		if (s0 == p)
			v = s1;
		return v;
	}
	return i1;
}

// flip test (src/simplifier.cpp:1200-1232) of rank k over every way the lower candidates still in
// play could move its neighbours: 0 undetermined, 1 passes in all, 2 fails in all.  The test
// reads collapse_remap at the corners of the triangles around k's source; a corner whose class
// published a lower source-role rank in the previous round may still move (every pending rank
// publishes every round), every other corner keeps the value k sees now.  A moving corner's
// alternatives are the value seen and the values its lower candidates' commits write, taken from
// the per-source candidate lists; a candidate counts unless it is known never to commit (skipped,
// flip-failed, rejected), so the set is never smaller than the true one, whatever stores of the
// same round the test sees.  An outcome that holds in every combination is final; a triangle
// without a moving corner decides alone.  Limited to kMaxAlternatives per corner and
// kMaxPendingSlots moving corners; beyond that the rank waits.
//
// The test is a chain of dependent gathers, so the adjacency is read in chunks of kFlipChunk
// entries, and a source list kSrcChunk ranks at a time, with the gathers of each level issued
// before any is used.

constexpr int kFlipChunk = 4;      // the deep test's adjacency chunk
constexpr int kCheapFlipChunk = 4; // the cheap test's
constexpr int kCheapWaves = 5;     // waves per SIMD the cheap round kernel is built for
constexpr int kSrcChunk = 8;
constexpr int kMaxAlternatives = 2; // per corner: its current value and one possible move
constexpr int kMaxPendingSlots = 4; // moving corners before the rank waits

// one adjacency entry with moving corners (pend2: bit h set when corner h moves): 0 mixed, 1 passes
// in all combinations, 2 fails in all, 3 too many alternatives
__device__ int flip_entry_pending(const RoundState& a, uint32_t k, const uint32_t nb[2], const uint32_t seen[2], uint32_t pend2, const V3& v0,
                                  const V3& v1, uint32_t i1)
{
	uint32_t alt[2][kMaxAlternatives];
	int na[2] = {1, 1};
	alt[0][0] = seen[0];
	alt[1][0] = seen[1];
	uint32_t tb[2], te[2];
#pragma unroll
	for (int q = 0; q < 2; ++q)
	{
		const bool scan = (pend2 >> q) & 1u;
		tb[q] = scan ? a.src_off[nb[q]] : 0;
		te[q] = scan ? a.src_off[nb[q] + 1] : 0;
	}
	bool more = true;
	for (uint32_t step = 0; more; step += kSrcChunk)
	{
		more = false;
		uint32_t r[2][kSrcChunk], val[2][kSrcChunk];
#pragma unroll
		for (int q = 0; q < 2; ++q)
#pragma unroll
			for (int u = 0; u < kSrcChunk; ++u)
			{
				const uint32_t t = tb[q] + step + u;
				const bool in = t < te[q];
				r[q][u] = in ? a.src_rank[t] : kNone;
				val[q][u] = in ? a.src_val[t] : 0;
			}
		uint8_t st[2][kSrcChunk];
#pragma unroll
		for (int q = 0; q < 2; ++q)
#pragma unroll
			for (int u = 0; u < kSrcChunk; ++u)
				st[q][u] = r[q][u] < k ? a.status[r[q][u]] : (uint8_t)kSkipped;
#pragma unroll
		for (int q = 0; q < 2; ++q)
		{
#pragma unroll
			for (int u = 0; u < kSrcChunk; ++u)
			{
				if (st[q][u] == kSkipped || st[q][u] == kRejected || st[q][u] == kFlipFail || val[q][u] == alt[q][0])
					continue;
				if (na[q] == kMaxAlternatives)
					return 3;
				alt[q][na[q]++] = val[q][u];
			}
			more |= tb[q] + step + kSrcChunk < te[q];
		}
	}
	bool any_flip = false, any_ok = false;
	for (int x = 0; x < na[0]; ++x)
		for (int y = 0; y < na[1]; ++y)
		{
			if (nb[0] == nb[1] && x != y)
				continue; // one vertex moves one way only
			uint32_t va = alt[0][x], vb = alt[1][y];
			if (va == i1 || vb == i1 || va == vb)
				any_ok = true;
			else if (triangle_flips(load_pos(a.pos, va), load_pos(a.pos, vb), v0, v1))
				any_flip = true;
			else
				any_ok = true;
		}
	return any_flip ? (any_ok ? 0 : 2) : 1;
}

// begin/end: rank k's source's adjacency range; v0, v1: the positions of its source and target.
// kDeep: examine the combinations of the moving corners' alternatives; otherwise a rank with a
// moving corner waits (the cheap test of the first rounds of a window, when most corners move)
template <bool kDeep, int kChunk = kDeep ? kFlipChunk : kCheapFlipChunk>
__device__ int flip_outcome(const RoundState& a, uint32_t i1, uint32_t k, uint32_t epoch, uint32_t begin, uint32_t end, const V3& v0, const V3& v1)
{
	const uint64_t* prev_pend = a.pend[(epoch - 1) & 1];
	const uint32_t prev_tag = ~(epoch - 1);
	int pending = 0;
	bool all_pass = true;
	for (uint32_t p0 = begin; p0 < end; p0 += kChunk)
	{
		const uint32_t n = end - p0 < (uint32_t)kChunk ? end - p0 : (uint32_t)kChunk;
		uint32_t nb[2 * kChunk];
#pragma unroll
		for (int j = 0; j < kChunk; ++j)
		{
			const uint2 e = (uint32_t)j < n ? a.adj[p0 + j] : make_uint2(i1, i1);
			nb[2 * j] = e.x;
			nb[2 * j + 1] = e.y;
		}
		uint32_t cm[2 * kChunk], lr[2 * kChunk];
		uint64_t ps[2 * kChunk];
#pragma unroll
		for (int q = 0; q < 2 * kChunk; ++q)
		{
			cm[q] = a.cremap[nb[q]];
			lr[q] = a.lock_rank[nb[q]];
			ps[q] = prev_pend[2 * nb[q]];
		}
		// collapse_remap as rank k sees it: a move committed by a higher rank (possible once
		// flip tests are decided early) reads as "not moved"
		uint32_t seen[2 * kChunk];
		uint32_t pend_mask = 0;
#pragma unroll
		for (int q = 0; q < 2 * kChunk; ++q)
		{
			seen[q] = (cm[q] != nb[q] && lr[q] > k) ? nb[q] : cm[q];
			const bool moving = (uint32_t)(q / 2) < n && (uint32_t)(ps[q] >> 32) == prev_tag && (uint32_t)ps[q] < k;
			pend_mask |= moving ? 1u << q : 0u;
		}
		pending += __builtin_popcount(pend_mask);
		if (kDeep && pending > kMaxPendingSlots)
			return 0;
		// entries without a moving corner have one scenario: the positions they need, all at once
		V3 pa[kChunk], pb[kChunk];
		uint32_t test_mask = 0;
#pragma unroll
		for (int j = 0; j < kChunk; ++j)
		{
			const uint32_t va = seen[2 * j], vb = seen[2 * j + 1];
			const bool single = (uint32_t)j < n && !((pend_mask >> (2 * j)) & 3u);
			const bool test = single && !(va == i1 || vb == i1 || va == vb);
			test_mask |= test ? 1u << j : 0u;
			pa[j] = test ? load_pos(a.pos, va) : V3{0.f, 0.f, 0.f};
			pb[j] = test ? load_pos(a.pos, vb) : V3{0.f, 0.f, 0.f};
		}
#pragma unroll
		for (int j = 0; j < kChunk; ++j)
			if (((test_mask >> j) & 1u) && triangle_flips(pa[j], pb[j], v0, v1))
				return 2;
		if (!kDeep)
		{
			all_pass &= pending == 0;
			continue;
		}
		// entries with a moving corner: every combination of the alternatives
		for (int j = 0; j < kChunk; ++j)
		{
			if ((uint32_t)j >= n || !((pend_mask >> (2 * j)) & 3u))
				continue;
			const uint32_t nbj[2] = {nb[2 * j], nb[2 * j + 1]};
			const uint32_t seenj[2] = {seen[2 * j], seen[2 * j + 1]};
			const int o = flip_entry_pending(a, k, nbj, seenj, (pend_mask >> (2 * j)) & 3u, v0, v1, i1);
			if (o == 3)
				return 0;
			if (o == 2)
				return 2;
			if (o == 0)
				all_pass = false;
		}
	}
	return all_pass ? 1 : 0;
}

// round `epoch`, one rank.  A rank at a locked endpoint is skipped.  A rank that was the lowest
// published at one of its endpoints in the previous round gets its flip test; one that was the
// lowest at both and whose flip test has passed commits now (the commit phase of the previous
// round: no lower rank is left at its endpoints, and ranks committing in one round share no
// endpoint and no neighbourhood in which one of them is still pending for the other, so they
// never touch each other's state).  The rest publish their rank at both endpoint classes and
// stay on the list (true).  Every gather the decision may need is issued before the first is
// used.
//
// Commits run concurrently with the other ranks' tests of the same round.  A rank's flip test
// only trusts corners without a lower source published in the previous round, and every rank
// that commits in this round published its source then, so no test reads a move still in
// flight.  A lock written in this round may be missed by a rank sharing that endpoint: that rank
// publishes once more and is skipped in the next round, before it can commit.
template <bool kDeep>
__device__ __forceinline__ bool round_x(const RoundState& a, uint4 w, uint32_t epoch)
{
	const uint32_t k = w.x, r0 = w.y, r1 = w.z;
	const ulonglong2* prev = reinterpret_cast<const ulonglong2*>(a.pend[(epoch - 1) & 1]);
	const uint8_t st = a.status[k];
	const uint32_t l0 = a.lock_rank[r0], l1 = a.lock_rank[r1];
	const ulonglong2 q0 = prev[r0], q1 = prev[r1];
	const uint64_t p0 = q0.x < q0.y ? q0.x : q0.y, p1 = q1.x < q1.y ? q1.x : q1.y;
	const uint32_t begin = a.adj_off[r0], end = a.adj_off[r0 + 1];
	const V3 v0 = load_pos(a.pos, r0), v1 = load_pos(a.pos, r1);
	if (st != kUndecided && st != kFlipPass)
		return false;
	if (l0 != kNone || l1 != kNone)
	{
		a.status[k] = kSkipped;
		return false;
	}
	// the flip test is only worth evaluating for a rank that was the lowest published at one of
	// its endpoints in the previous round (others cannot commit yet)
	const uint64_t was = round_key(epoch - 1, k);
	const bool low0 = p0 == was, low1 = p1 == was;
	uint8_t now = st;
	if (st == kUndecided && (low0 || low1))
	{
		int o = flip_outcome<kDeep>(a, r1, k, epoch, begin, end, v0, v1);
		if (o == 2)
		{
			a.status[k] = kFlipFail;
			return false;
		}
		if (o == 1)
			now = kFlipPass;
	}
	if (now == kFlipPass && low0 && low1)
	{
		commit_collapse(a.cremap, a.r_v0[k], a.r_v1[k], a.kind, a.wedge, a.remap, a.loop, a.loopback);
		a.lock_rank[r0] = k;
		a.lock_rank[r1] = k;
		a.status[k] = kCommitted;
		return false;
	}
	if (now != st)
		a.status[k] = now;
	// two atomics: the source-role minimum at r0 and the target-role one at r1
	uint64_t key = round_key(epoch, k);
	unsigned long long* wr = (unsigned long long*)a.pend[epoch & 1];
	atomicMin(&wr[2 * (size_t)r0], (unsigned long long)key);
	atomicMin(&wr[2 * (size_t)r1 + 1], (unsigned long long)key);
	return true;
}

__device__ __forceinline__ void append_wave(bool keep, uint4 w, uint4* out_list, uint32_t* out_count)
{
	const int lane = __lane_id();
	uint64_t mask = __ballot(keep);
	if (mask)
	{
		uint32_t slot0 = 0;
		int leader = __ffsll((long long)mask) - 1;
		if (lane == leader)
			slot0 = atomicAdd(out_count, (uint32_t)__popcll(mask));
		slot0 = __shfl(slot0, leader, 64);
		if (keep)
			out_list[slot0 + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull))] = w;
	}
}

// head round.  Round r reads list (r & 1) / counters (r % 3), appends the survivors to list
// ((r + 1) & 1) / counters ((r + 1) % 3) and clears counters ((r + 2) % 3).  A rank stays in its
// segment, at the segment's offset in both lists.  Workgroups are dealt to the XCDs in turn, so
// workgroup i works on segment i mod kListSegments: each XCD's L2 then holds the class state of
// one segment's ranks rather than of all of them.  The grid is a multiple of kListSegments; one
// rank per thread when it is large enough, and the grid-stride loop covers any segment length.
// kDeep: this round examines the alternatives of moving corners (the window's per-source lists
// are built).  Two kernels rather than a flag, so that the cheap rounds are not held to the deep
// test's register count.
template <bool kDeep>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(kDeep ? 4 : kCheapWaves))) k_round_x(RoundState a, uint32_t head_index, uint32_t rel)
{
	CallScalars* sc = a.sc;
	// rounds are numbered from a base on the device (the tail leaves it past its last round), so
	// rounds can follow a tail or a device-side window extension without a read-back; `rel` counts
	// the head rounds launched since the base was set
	const uint32_t round = sc->round + rel, epoch = sc->epoch + 1 + rel;
	const uint32_t seg = blockIdx.x % kListSegments;
	const uint32_t n = sc->list_count[round % 3][seg];
	const uint32_t off = sc->list_off[seg];
	if (blockIdx.x == 0)
	{
		if (threadIdx.x < kListSegments)
			sc->list_count[(round + 2) % 3][threadIdx.x] = 0;
		if (threadIdx.x == 0)
		{
			uint32_t total = 0;
			for (int q = 0; q < kListSegments; ++q)
				total += sc->list_count[round % 3][q];
			if (head_index < kHeadSizeSlots)
				sc->head_sizes[head_index] = total;
			sc->x_visits += total;
			sc->x_launches += 1;
		}
	}
	const uint4* in_list = a.list[round & 1] + off;
	uint4* out_list = a.list[(round + 1) & 1] + off;
	uint32_t* out_count = &sc->list_count[(round + 1) % 3][seg];
	const uint32_t stride = (gridDim.x / kListSegments) * kThreads;
	for (uint32_t base = (blockIdx.x / kListSegments) * kThreads; base < n; base += stride)
	{
		uint32_t idx = base + threadIdx.x;
		uint4 w = idx < n ? in_list[idx] : make_uint4(0, 0, 0, 0);
		bool keep = idx < n && round_x<kDeep>(a, w, epoch);
		append_wave(keep, w, out_list, out_count);
	}
}

constexpr int kTailThreads = 1024;

// tail: one workgroup runs the remaining rounds back to back (phases separated by barriers)
// until the list is empty.  The lowest undecided rank decides in every round, so a list of n
// ranks empties within n + 1 rounds; max_rounds bounds the loop regardless.  The first round
// reads the segments of the list; the survivors are appended as one list from offset 0 (segment
// 0, the others empty), which the later rounds read.
__global__ void __launch_bounds__(kTailThreads) k_round_tail(RoundState a, uint32_t rel, uint32_t max_rounds, bool deep)
{
	__shared__ uint32_t s_n, s_out;
	__shared__ uint32_t s_pre[kListSegments + 1], s_off[kListSegments];
	CallScalars* sc = a.sc;
	const uint32_t round0 = sc->round + rel, epoch0 = sc->epoch + 1 + rel;
	__syncthreads(); // every thread has read the base before thread 0 moves it below
	if (threadIdx.x == 0)
	{
		uint32_t total = 0;
		for (int q = 0; q < kListSegments; ++q)
		{
			s_pre[q] = total;
			s_off[q] = sc->list_off[q];
			total += sc->list_count[round0 % 3][q];
		}
		s_pre[kListSegments] = total;
		s_n = total;
		s_out = 0;
	}
	__syncthreads();

	uint32_t round = round0, epoch = epoch0, done = 0;
	for (;;)
	{
		uint32_t n = s_n;
		if (n == 0 || done >= max_rounds)
			break;
		const uint4* in_list = a.list[round & 1];
		uint4* out_list = a.list[(round + 1) & 1];
		for (uint32_t base = 0; base < n; base += kTailThreads)
		{
			uint32_t idx = base + threadIdx.x;
			uint4 w = make_uint4(0, 0, 0, 0);
			if (idx < n)
			{
				int q = 0;
				while (idx >= s_pre[q + 1]) // stops below kListSegments: idx < s_pre[kListSegments] = n
					++q;
				w = in_list[s_off[q] + (idx - s_pre[q])];
			}
			bool keep = idx < n && (deep ? round_x<true>(a, w, epoch) : round_x<false>(a, w, epoch));
			append_wave(keep, w, out_list, &s_out);
		}
		__syncthreads();
		if (threadIdx.x == 0)
		{
			s_n = s_out;
			s_out = 0;
			s_pre[0] = 0;
			for (int q = 1; q <= kListSegments; ++q)
				s_pre[q] = s_n;
			for (int q = 0; q < kListSegments; ++q)
				s_off[q] = 0;
		}
		round++;
		epoch++;
		done++;
		__syncthreads();
	}

	if (threadIdx.x == 0)
	{
		for (int q = 0; q < kListSegments; ++q)
		{
			sc->list_count[round % 3][q] = q == 0 ? s_n : 0u;
			sc->list_count[(round + 1) % 3][q] = 0;
			sc->list_count[(round + 2) % 3][q] = 0;
			sc->list_off[q] = 0;
			sc->window_count[q] = 0; // the next window open counts from zero
		}
		sc->tail_rounds += done;
		sc->round = round;
		sc->epoch = epoch - 1;
		if (s_n)
			sc->work[3] = 1; // round guard hit: the host reports an error
	}
}

__global__ void k_perform_init_scalars(CallScalars* sc)
{
	for (int q = 0; q < kListSegments; ++q)
	{
		sc->list_count[0][q] = 0;
		sc->list_count[1][q] = 0;
		sc->list_count[2][q] = 0;
		sc->list_off[q] = 0;
		sc->window_count[q] = 0;
	}
	sc->stop_reason = 0;
	sc->tail_rounds = 0;
	sc->round = 0;
	sc->extended = 0;
}

// candidate ranks grouped by source class (order inside a group is irrelevant): one atomic pass
// counts the ranks per source class and hands each its slot in the class's group; after the scan
// the fill places the ranks without atomics
__global__ void k_src_count(const uint32_t* r_r0, size_t count, uint32_t* counts, uint32_t* slot)
{
	size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (k < count)
		slot[k] = atomicAdd(&counts[r_r0[k]], 1u);
}

__global__ void k_src_fill(RoundState a, const uint32_t* r_r0, size_t count, const uint32_t* slot, uint32_t* src_rank, uint32_t* src_val)
{
	size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= count)
		return;
	uint32_t r0 = r_r0[k];
	uint32_t at = a.src_off[r0] + slot[k];
	src_rank[at] = (uint32_t)k;
	src_val[at] = commit_value_at(a, (uint32_t)k, r0);
}

static RoundState round_state(Mesh& m)
{
	RoundState a;
	a.r_v0 = m.r_v0;
	a.r_v1 = m.r_v1;
	a.adj_off = m.adj_off;
	a.adj = m.adj;
	a.pos = m.pos;
	a.kind = m.kind;
	a.wedge = m.wedge;
	a.remap = m.remap;
	a.loop = m.loop[m.loop_cur];
	a.loopback = m.loopback[m.loop_cur];
	a.src_off = m.src_off;
	a.src_rank = m.src_rank;
	a.src_val = m.src_val;
	a.cremap = m.cremap;
	a.lock_rank = m.lock_rank;
	a.status = m.status;
	for (int b = 0; b < 2; ++b)
	{
		a.pend[b] = m.pend[b];
		a.list[b] = m.list[b];
	}
	a.sc = m.sc;
	a.seg_mul = m.vertex_count ? ((uint64_t)kListSegments << 32) / m.vertex_count : 0;
	return a;
}

void launch_perform_init(hipStream_t s, Mesh& m, size_t cand_count, bool source_lists)
{
	size_t n = m.vertex_count;
	iota_u32(s, m.cremap, n);
	fill_u32(s, m.lock_rank, kNone, n);
	k_perform_init_scalars<<<1, 1, 0, s>>>(m.sc);
	if (source_lists)
		launch_source_lists(s, m, cand_count);
}

// the per-source candidate lists deep rounds read (they depend on the ranks alone, so they may
// also be built between the rounds of a window)
void launch_source_lists(hipStream_t s, Mesh& m, size_t cand_count)
{
	size_t n = m.vertex_count;
	fill_u32(s, m.src_off, 0, n + 1);
	// the rank slots go to m.keys, free once the ranked copies are gathered
	if (cand_count)
		k_src_count<<<grid_for(cand_count), kThreads, 0, s>>>(m.r_r0, cand_count, m.src_off, m.keys);
	scan_exclusive(s, m.src_off, m.src_off, n, nullptr, m.src_off + n, m.temp);
	if (cand_count)
		k_src_fill<<<grid_for(cand_count), kThreads, 0, s>>>(round_state(m), m.r_r0, cand_count, m.keys, m.src_rank, m.src_val);
}

// ranks [lo, hi) become the list of the next round (sc->round), in segments by source class.
// With `if_no_stop` the window opens only when the stop scan of [0, lo) found no break
// (sc->stop == lo) and the list goes on: the rest of the list is then decided in the same pass
// without a read-back (sc->extended = 1); otherwise the list stays empty and the rounds that
// follow do nothing.  Two kernels: the count of each segment (into sc->window_count, zero since
// the previous window's tail), then the placement, each segment at the sum of the counts before it.
__device__ __forceinline__ bool window_opens(const CallScalars* sc, uint32_t lo, uint32_t hi, bool if_no_stop)
{
	return !if_no_stop || (sc->stop >= lo && hi > lo);
}

__global__ void __launch_bounds__(kThreads) k_window_count(RoundState a, uint32_t lo, uint32_t hi, const uint32_t* r_r0, bool if_no_stop)
{
	__shared__ uint32_t s_cnt[kListSegments];
	CallScalars* sc = a.sc;
	const uint32_t round = sc->round;
	const bool open = window_opens(sc, lo, hi, if_no_stop);
	if (threadIdx.x < kListSegments)
	{
		s_cnt[threadIdx.x] = 0;
		if (blockIdx.x == 0)
		{
			sc->list_count[round % 3][threadIdx.x] = 0; // filled by the placement
			sc->list_count[(round + 1) % 3][threadIdx.x] = 0;
		}
	}
	__syncthreads();
	const uint32_t k = lo + blockIdx.x * blockDim.x + threadIdx.x;
	if (open && k < hi)
		atomicAdd(&s_cnt[list_segment(a, r_r0[k])], 1u);
	__syncthreads();
	if (threadIdx.x < kListSegments && s_cnt[threadIdx.x])
		atomicAdd(&sc->window_count[threadIdx.x], s_cnt[threadIdx.x]);
}

__global__ void __launch_bounds__(kThreads) k_window_open(RoundState a, uint32_t lo, uint32_t hi, const uint32_t* r_r0, const uint32_t* r_r1, bool if_no_stop)
{
	__shared__ uint32_t s_off[kListSegments], s_cnt[kListSegments], s_base[kListSegments];
	CallScalars* sc = a.sc;
	const uint32_t round = sc->round;
	const bool open = window_opens(sc, lo, hi, if_no_stop);
	if (threadIdx.x < kListSegments)
	{
		uint32_t off = 0;
		for (uint32_t q = 0; q < threadIdx.x; ++q)
			off += sc->window_count[q];
		s_off[threadIdx.x] = off;
		s_cnt[threadIdx.x] = 0;
		if (blockIdx.x == 0)
			sc->list_off[threadIdx.x] = off;
	}
	if (if_no_stop && blockIdx.x == 0 && threadIdx.x == 0)
		sc->extended = open ? 1u : 0u;
	__syncthreads();
	const uint32_t k = lo + blockIdx.x * blockDim.x + threadIdx.x;
	const bool in = open && k < hi;
	const uint32_t r0 = in ? r_r0[k] : 0;
	const uint32_t seg = list_segment(a, r0);
	uint32_t local = 0;
	if (in)
		local = atomicAdd(&s_cnt[seg], 1u);
	__syncthreads();
	if (threadIdx.x < kListSegments && s_cnt[threadIdx.x])
		s_base[threadIdx.x] = atomicAdd(&sc->list_count[round % 3][threadIdx.x], s_cnt[threadIdx.x]);
	__syncthreads();
	if (in)
		a.list[round & 1][s_off[seg] + s_base[seg] + local] = make_uint4(k, r0, r_r1[k], 0);
}

void launch_window_extend(hipStream_t s, Mesh& m, size_t lo, size_t hi, bool if_no_stop)
{
	const unsigned int grid = grid_for(hi > lo ? hi - lo : 1);
	const RoundState a = round_state(m);
	k_window_count<<<grid, kThreads, 0, s>>>(a, (uint32_t)lo, (uint32_t)hi, m.r_r0, if_no_stop);
	k_window_open<<<grid, kThreads, 0, s>>>(a, (uint32_t)lo, (uint32_t)hi, m.r_r0, m.r_r1, if_no_stop);
}

// list_estimate sizes the grid (one rank per thread); the kernels stay correct for any length
void launch_round(hipStream_t s, Mesh& m, size_t list_estimate, int head_index, int rel, bool deep)
{
	// a multiple of kListSegments workgroups, at least one per segment
	unsigned int blocks = grid_for(list_estimate);
	blocks = (blocks + kListSegments - 1) / kListSegments * kListSegments;
	blocks = blocks < kListSegments ? kListSegments : (blocks > 8192 ? 8192 : blocks);
	RoundState a = round_state(m);
	KernelTimer* t = m.timer;
	bool timed = t && t->count < t->capacity;
	if (timed)
		HIP_CHECK(hipEventRecord(t->begin[t->count], s));
	if (deep)
		k_round_x<true><<<blocks, kThreads, 0, s>>>(a, (uint32_t)head_index, (uint32_t)rel);
	else
		k_round_x<false><<<blocks, kThreads, 0, s>>>(a, (uint32_t)head_index, (uint32_t)rel);
	if (timed)
		HIP_CHECK(hipEventRecord(t->end[t->count++], s));
}

void launch_round_tail(hipStream_t s, Mesh& m, int rel, uint32_t max_rounds, bool deep)
{
	k_round_tail<<<1, kTailThreads, 0, s>>>(round_state(m), (uint32_t)rel, max_rounds, deep);
}

// a failed flip test is a rejection unless a lower commit locked an endpoint first
__global__ void k_resolve_flip_fail(uint8_t* status, const uint32_t* r_r0, const uint32_t* r_r1, const uint32_t* lock_rank, uint32_t lo, uint32_t hi)
{
	uint32_t k = lo + blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= hi || status[k] != kFlipFail)
		return;
	uint32_t l0 = lock_rank[r_r0[k]], l1 = lock_rank[r_r1[k]];
	status[k] = ((l0 != kNone && l0 < k) || (l1 != kNone && l1 < k)) ? kSkipped : kRejected;
}

void launch_resolve(hipStream_t s, Mesh& m, size_t lo, size_t hi)
{
	if (hi > lo)
		k_resolve_flip_fail<<<grid_for(hi - lo), kThreads, 0, s>>>(m.status, m.r_r0, m.r_r1, m.lock_rank, (uint32_t)lo, (uint32_t)hi);
}

// ------------------------------------------------------------------------------------------
// stop scan: the first rank where the sequential loop breaks (src/simplifier.cpp:1517-1539)
//   S1 err > error_limit
//   S2 tri_collapses >= goal
//   S3 err > 1.5 * err[goal/2 + rejects] && err > result_error && tri_collapses > goal/6
// with tri_collapses, rejects and result_error the exclusive prefix sums / max of the
// decided ranks.  Tiles of 4096 ranks: reduce, scan the tile aggregates, then test.

struct StopAgg
{
	uint32_t tri, rej, err;
};

__device__ __forceinline__ StopAgg stop_item(uint8_t st, uint8_t k0, float err)
{
	StopAgg a;
	a.tri = st == kCommitted ? (k0 == kBorder ? 1u : 2u) : 0u;
	a.rej = st == kRejected ? 1u : 0u;
	// result_error keeps a NaN error out (its `<` test fails); errors are >= 0 otherwise
	a.err = (st == kCommitted && err == err) ? __float_as_uint(err) : 0u;
	return a;
}

template <bool Max>
__device__ __forceinline__ uint32_t wave_incl(uint32_t v)
{
	const int lane = __lane_id();
#pragma unroll
	for (int o = 1; o < 64; o <<= 1)
	{
		uint32_t t = __shfl_up(v, o, 64);
		if (lane >= o)
			v = Max ? (t > v ? t : v) : v + t;
	}
	return v;
}

// exclusive block scan of three values (sum, sum, max) for 256 threads
__device__ __forceinline__ StopAgg block_scan3(StopAgg v, StopAgg& total)
{
	__shared__ uint32_t lds[3][4];
	uint32_t it = wave_incl<false>(v.tri);
	uint32_t ir = wave_incl<false>(v.rej);
	uint32_t ie = wave_incl<true>(v.err);
	const int wave = threadIdx.x >> 6;
	if (__lane_id() == 63)
	{
		lds[0][wave] = it;
		lds[1][wave] = ir;
		lds[2][wave] = ie;
	}
	__syncthreads();
	uint32_t bt = 0, br = 0, be = 0, at = 0, ar = 0, ae = 0;
#pragma unroll
	for (int w = 0; w < 4; ++w)
	{
		uint32_t t = lds[0][w], r = lds[1][w], e = lds[2][w];
		if (w < wave)
		{
			bt += t;
			br += r;
			be = e > be ? e : be;
		}
		at += t;
		ar += r;
		ae = e > ae ? e : ae;
	}
	__syncthreads();
	total = StopAgg{at, ar, ae};
	// exclusive: inclusive minus own (sums); for the max, the inclusive value of the previous lane
	uint32_t pe = __shfl_up(ie, 1, 64);
	if (__lane_id() == 0)
		pe = 0;
	StopAgg ex;
	ex.tri = bt + it - v.tri;
	ex.rej = br + ir - v.rej;
	ex.err = be > pe ? be : pe;
	return ex;
}

__global__ void __launch_bounds__(kThreads) k_stop_reduce(const uint8_t* status, const uint32_t* r_v0, const uint8_t* kind, const float* r_err,
                                                          size_t count, uint32_t* tiles, const CallScalars* sc, bool conditional)
{
	if (conditional && !sc->extended)
		return;
	size_t base = (size_t)blockIdx.x * kScanTile;
	StopAgg acc{0, 0, 0};
	for (int j = 0; j < kScanItems; ++j)
	{
		size_t k = base + (size_t)j * kThreads + threadIdx.x;
		if (k < count)
		{
			StopAgg a = stop_item(status[k], kind[r_v0[k]], r_err[k]);
			acc.tri += a.tri;
			acc.rej += a.rej;
			acc.err = a.err > acc.err ? a.err : acc.err;
		}
	}
	StopAgg total;
	block_scan3(acc, total);
	if (threadIdx.x == 0)
	{
		tiles[blockIdx.x * 3 + 0] = total.tri;
		tiles[blockIdx.x * 3 + 1] = total.rej;
		tiles[blockIdx.x * 3 + 2] = total.err;
	}
}

// single block: exclusive prefix of the tile aggregates; the max starts at the incoming result_error
__global__ void __launch_bounds__(kThreads) k_stop_tiles(uint32_t* tiles, uint32_t ntiles, const CallScalars* sc, bool conditional)
{
	if (conditional && !sc->extended)
		return;
	StopAgg carry{0, 0, __float_as_uint(sc->result_error)};
	for (uint32_t b = 0; b < ntiles; b += kThreads)
	{
		uint32_t i = b + threadIdx.x;
		StopAgg v{0, 0, 0};
		if (i < ntiles)
			v = StopAgg{tiles[i * 3 + 0], tiles[i * 3 + 1], tiles[i * 3 + 2]};
		StopAgg total;
		StopAgg ex = block_scan3(v, total);
		if (i < ntiles)
		{
			tiles[i * 3 + 0] = carry.tri + ex.tri;
			tiles[i * 3 + 1] = carry.rej + ex.rej;
			uint32_t e = ex.err > carry.err ? ex.err : carry.err;
			tiles[i * 3 + 2] = e;
		}
		carry.tri += total.tri;
		carry.rej += total.rej;
		carry.err = total.err > carry.err ? total.err : carry.err;
	}
}

__global__ void __launch_bounds__(kThreads) k_stop_find(const uint8_t* status, const uint32_t* r_v0, const uint8_t* kind, const float* r_err,
                                                        size_t count, size_t cand_count, const uint32_t* tiles, uint32_t goal, float error_limit,
                                                        CallScalars* sc, bool conditional)
{
	if (conditional && !sc->extended)
		return;
	size_t base = (size_t)blockIdx.x * kScanTile;
	StopAgg carry{tiles[blockIdx.x * 3 + 0], tiles[blockIdx.x * 3 + 1], tiles[blockIdx.x * 3 + 2]};
	const uint32_t half = goal / 2, sixth = goal / 6;
	// ranks grow with the lane and the item, so a wavefront's first stopping rank is its minimum:
	// one atomic per wavefront (stopping ranks form long runs; same-address atomics serialise)
	bool found = false;
	for (int j = 0; j < kScanItems; ++j)
	{
		size_t k = base + (size_t)j * kThreads + threadIdx.x;
		StopAgg v{0, 0, 0};
		float err = 0.f;
		if (k < count)
		{
			err = r_err[k];
			v = stop_item(status[k], kind[r_v0[k]], err);
		}
		StopAgg total;
		StopAgg ex = block_scan3(v, total);
		bool stop = false;
		if (k < count)
		{
			uint32_t tri = carry.tri + ex.tri;
			uint32_t rej = carry.rej + ex.rej;
			uint32_t eb = ex.err > carry.err ? ex.err : carry.err;
			float result_error = __uint_as_float(eb);
			size_t eg = (size_t)half + rej;
			float error_goal = eg < cand_count ? 1.5f * r_err[eg] : FLT_MAX;
			stop = err > error_limit || tri >= goal || (err > error_goal && err > result_error && tri > sixth);
		}
		const uint64_t stops = __ballot(stop);
		if (stops && !found)
		{
			found = true;
			if (__lane_id() == (uint32_t)__builtin_ctzll(stops))
				atomicMin(&sc->stop, (uint32_t)k);
		}
		carry.tri += total.tri;
		carry.rej += total.rej;
		carry.err = total.err > carry.err ? total.err : carry.err;
	}
}

// totals below the stop rank (one atomic per workgroup and counter)
__global__ void __launch_bounds__(kThreads) k_stop_totals(const uint8_t* status, const uint32_t* r_v0, const uint8_t* kind, const float* r_err,
                                                          CallScalars* sc, bool conditional)
{
	if (conditional && !sc->extended)
		return;
	__shared__ uint32_t red[5][4];
	uint32_t stop = sc->stop;
	uint32_t v[5] = {0, 0, 0, 0, 0}; // committed, skipped, rejected, tri, max error bits
	for (uint32_t k = blockIdx.x * kThreads + threadIdx.x; k < stop; k += gridDim.x * kThreads)
	{
		uint8_t st = status[k];
		StopAgg a = stop_item(st, kind[r_v0[k]], r_err[k]);
		v[0] += st == kCommitted;
		v[1] += st == kSkipped;
		v[2] += st == kRejected;
		v[3] += a.tri;
		v[4] = a.err > v[4] ? a.err : v[4];
	}
#pragma unroll
	for (int o = 32; o > 0; o >>= 1)
	{
#pragma unroll
		for (int q = 0; q < 4; ++q)
			v[q] += __shfl_xor(v[q], o, 64);
		uint32_t t = __shfl_xor(v[4], o, 64);
		v[4] = t > v[4] ? t : v[4];
	}
	const int wave = threadIdx.x >> 6;
	if (__lane_id() == 0)
		for (int q = 0; q < 5; ++q)
			red[q][wave] = v[q];
	__syncthreads();
	if (threadIdx.x == 0)
	{
		uint32_t t[5] = {0, 0, 0, 0, 0};
		for (int w = 0; w < kThreads / 64; ++w)
		{
			for (int q = 0; q < 4; ++q)
				t[q] += red[q][w];
			t[4] = red[4][w] > t[4] ? red[4][w] : t[4];
		}
		if (t[0])
			atomicAdd(&sc->committed, t[0]);
		if (t[1])
			atomicAdd(&sc->skipped, t[1]);
		if (t[2])
			atomicAdd(&sc->rejected, t[2]);
		if (t[3])
			atomicAdd(&sc->tri_collapses, t[3]);
		if (t[4])
			atomicMax(&sc->work[1], t[4]);
	}
}

__global__ void k_stop_finish(const float* r_err, uint32_t count, uint32_t goal, float error_limit, CallScalars* sc)
{
	uint32_t stop = sc->stop;
	float prev = sc->result_error;
	float best = __uint_as_float(sc->work[1]);
	sc->result_error = prev < best ? best : prev;
	uint32_t reason = 0;
	if (stop < count)
	{
		float err = r_err[stop];
		if (err > error_limit)
			reason = 1;
		else if (sc->tri_collapses >= goal)
			reason = 2;
		else
			reason = 3;
	}
	sc->stop_reason = reason;
}

// undo commits at or past the stop rank: a commit wrote only its source class's wedges
__global__ void k_revert(const uint8_t* status, const uint32_t* r_r0, const uint32_t* wedge, size_t count, const CallScalars* sc, uint32_t* cremap)
{
	size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= count || k < sc->stop || status[k] != kCommitted)
		return;
	uint32_t r = r_r0[k];
	uint32_t v = r;
	do
	{
		cremap[v] = v;
		v = wedge[v];
	} while (v != r);
}

__global__ void k_stop_reset(CallScalars* sc, uint32_t window, bool conditional)
{
	if (conditional && !sc->extended)
		return;
	sc->stop = window;
	sc->committed = 0;
	sc->skipped = 0;
	sc->rejected = 0;
	sc->tri_collapses = 0;
	sc->work[1] = 0;
}

// repeatable: reads the decisions of ranks [0, window) and writes only the stop rank (window
// if no rank in it breaks the loop) and the totals below it.  `conditional`: only when the window
// was extended on the device (sc->extended), else every kernel leaves at once
void launch_stop_scan(hipStream_t s, Mesh& m, size_t window, size_t cand_count, uint32_t goal, float error_limit, bool conditional)
{
	k_stop_reset<<<1, 1, 0, s>>>(m.sc, (uint32_t)window, conditional);
	if (!window)
		return;
	unsigned int tiles = div_up(window, kScanTile);
	k_stop_reduce<<<tiles, kThreads, 0, s>>>(m.status, m.r_v0, m.kind, m.r_err, window, m.stop_tile, m.sc, conditional);
	k_stop_tiles<<<1, kThreads, 0, s>>>(m.stop_tile, tiles, m.sc, conditional);
	k_stop_find<<<tiles, kThreads, 0, s>>>(m.status, m.r_v0, m.kind, m.r_err, window, cand_count, m.stop_tile, goal, error_limit, m.sc, conditional);
	unsigned int blocks = div_up(window, kThreads * 4);
	blocks = blocks < 1 ? 1 : (blocks > 512 ? 512 : blocks);
	k_stop_totals<<<blocks, kThreads, 0, s>>>(m.status, m.r_v0, m.kind, m.r_err, m.sc, conditional);
}

// once every rank of the window is decided: advance result_error and undo the commits at or
// past the stop rank
void launch_stop_commit(hipStream_t s, Mesh& m, size_t window, uint32_t goal, float error_limit)
{
	k_stop_finish<<<1, 1, 0, s>>>(m.r_err, (uint32_t)window, goal, error_limit, m.sc);
	if (window)
		k_revert<<<grid_for(window), kThreads, 0, s>>>(m.status, m.r_r0, m.wedge, window, m.sc, m.cremap);
}

// ------------------------------------------------------------------------------------------
// updateQuadrics (src/simplifier.cpp:1638-1673), one thread per commit below the stop.
// The target class receives exactly one positional add (its primary is locked); attribute
// quadrics move wedge by wedge in ascending wedge order, as the reference's vertex loop does.

__device__ __forceinline__ uint32_t next_wedge_above(const uint32_t* wedge, uint32_t r, uint32_t after)
{
	// smallest member of r's ring greater than `after` (kNone if none)
	uint32_t best = kNone;
	uint32_t v = r;
	do
	{
		if (v > after && v < best)
			best = v;
		v = wedge[v];
	} while (v != r);
	return best;
}

// one committed collapse of k_update_quadrics; returns the bits of its non-NaN positional error
// (non-negative floats order like their bits), 0 without attributes
__device__ __forceinline__ uint32_t update_one(size_t k, const uint32_t* r_r0, const uint32_t* r_r1, const uint32_t* cremap, const uint32_t* wedge,
                                               const float4* pos, int attr_count, Quad* vq, Grad* vg, Quad* aq, Grad* ag)
{
	uint32_t r0 = r_r0[k], r1 = r_r1[k];

	Quad Q0 = load_quad(vq, r0);
	Quad Q1 = load_quad(vq, r1);
	quad_add(Q1, Q0);
	store_quad(vq, r1, Q1);
	if (vg)
	{
		// the class is a target of one collapse per pass at most: no other update of vg[r1]
		Grad G1 = vg[r1];
		grad_add(G1, vg[r0]);
		vg[r1] = G1;
	}

	if (attr_count)
	{
		const int A = attr_count;
		// r0 is the lowest index of its class, so ascending order starts there
		for (uint32_t v = r0; v != kNone; v = next_wedge_above(wedge, r0, v))
		{
			uint32_t t = cremap[v];
			if (t == v)
				continue;
			Quad Qt = load_quad(aq, t);
			quad_add(Qt, load_quad(aq, v));
			store_quad(aq, t, Qt);
			for (int a = 0; a < A; ++a)
			{
				Grad g = ag[(size_t)t * A + a];
				grad_add(g, ag[(size_t)v * A + a]);
				ag[(size_t)t * A + a] = g;
			}
		}

		float derr = quad_error(Q0, load_pos(pos, r1));
		if (derr == derr)
			return __float_as_uint(derr);
	}
	return 0;
}

__global__ void k_update_quadrics(const uint8_t* status, const uint32_t* r_r0, const uint32_t* r_r1, const CallScalars* sc_in, const uint32_t* cremap,
                                  const uint32_t* wedge, const float4* pos, int attr_count, Quad* vq, Grad* vg, Quad* aq, Grad* ag, CallScalars* sc)
{
	size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	const bool commit = k < sc_in->stop && status[k] == kCommitted;
	// the largest positional error of the pass's commits (attribute path): one atomic per
	// wavefront, so every lane stays to the reduction
	uint32_t err_bits = 0;
	if (commit)
		err_bits = update_one(k, r_r0, r_r1, cremap, wedge, pos, attr_count, vq, vg, aq, ag);
	if (attr_count)
	{
#pragma unroll
		for (int o = 32; o > 0; o >>= 1)
		{
			uint32_t t = __shfl_xor(err_bits, o, 64);
			err_bits = t > err_bits ? t : err_bits;
		}
		if (__lane_id() == 0 && err_bits)
			atomicMax(&sc->vertex_error_bits, err_bits);
	}
}

// remapEdgeLoops, src/simplifier.cpp:1845-1864.  collapse_remap is idempotent within a pass
// (targets are locked, never sources), so reading loop[l] before or after its own update
// gives the same result and the in-place loop becomes a double-buffered map.
__global__ void k_remap_loops(const uint32_t* cremap, size_t n, const uint32_t* loop_in, uint32_t* loop_out, const uint32_t* back_in, uint32_t* back_out)
{
	size_t ii = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (ii >= n)
		return;
	uint32_t i = (uint32_t)ii;
	const uint32_t* in[2] = {loop_in, back_in};
	uint32_t* out[2] = {loop_out, back_out};
#pragma unroll
	for (int q = 0; q < 2; ++q)
	{
		uint32_t l = in[q][i];
		uint32_t res = kNone;
		if (l != kNone)
		{
			uint32_t r = cremap[l];
			if (i == r)
			{
				uint32_t ll = in[q][l];
				res = ll != kNone ? cremap[ll] : kNone;
			}
			else
				res = r;
		}
		out[q][i] = res;
	}
}

__global__ void k_vertex_error_from_result(CallScalars* sc)
{
	sc->vertex_error = sc->result_error;
}

__global__ void k_vertex_error_from_bits(CallScalars* sc)
{
	float v = __uint_as_float(sc->vertex_error_bits);
	float prev = sc->vertex_error;
	sc->vertex_error = prev < v ? v : prev;
}

void launch_apply(hipStream_t s, Mesh& m, size_t cand_count)
{
	if (cand_count)
		k_update_quadrics<<<grid_for(cand_count), kThreads, 0, s>>>(m.status, m.r_r0, m.r_r1, m.sc, m.cremap, m.wedge, m.pos, (int)m.attr_count,
		                                                           m.vq, m.vg, m.aq, m.ag, m.sc);
	if (m.attr_count)
		k_vertex_error_from_bits<<<1, 1, 0, s>>>(m.sc);
	else
		k_vertex_error_from_result<<<1, 1, 0, s>>>(m.sc);

	int cur = m.loop_cur;
	if (m.vertex_count)
		k_remap_loops<<<grid_for(m.vertex_count), kThreads, 0, s>>>(m.cremap, m.vertex_count, m.loop[cur], m.loop[cur ^ 1], m.loopback[cur],
		                                                          m.loopback[cur ^ 1]);
	m.loop_cur = cur ^ 1;
}

// ------------------------------------------------------------------------------------------
// remapIndexBuffer, src/simplifier.cpp:1811-1843: apply collapse_remap, drop triangles whose
// primaries are no longer distinct, keep the survivors' order

__device__ __forceinline__ bool remap_triangle(const uint32_t* ib, size_t t, const uint32_t* cremap, const uint32_t* remap, uint32_t v[3])
{
	v[0] = cremap[ib[t * 3 + 0]];
	v[1] = cremap[ib[t * 3 + 1]];
	v[2] = cremap[ib[t * 3 + 2]];
	uint32_t r0 = remap[v[0]], r1 = remap[v[1]], r2 = remap[v[2]];
	return r0 != r1 && r0 != r2 && r1 != r2;
}

__global__ void k_tri_keep(const uint32_t* ib, size_t tris, const uint32_t* cremap, const uint32_t* remap, uint32_t* keep)
{
	size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= tris)
		return;
	uint32_t v[3];
	keep[t] = remap_triangle(ib, t, cremap, remap, v);
}

__global__ void k_tri_write(const uint32_t* ib, size_t tris, const uint32_t* cremap, const uint32_t* remap, const uint32_t* offsets, uint32_t* out)
{
	size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= tris)
		return;
	uint32_t v[3];
	if (remap_triangle(ib, t, cremap, remap, v))
	{
		size_t o = (size_t)offsets[t] * 3;
		out[o + 0] = v[0];
		out[o + 1] = v[1];
		out[o + 2] = v[2];
	}
}

void launch_remap_triangles(hipStream_t s, Mesh& m, size_t tri_count)
{
	HIP_CHECK(hipMemsetAsync(&m.sc->next_triangles, 0, sizeof(uint32_t), s));
	if (!tri_count)
		return;
	const uint32_t* in = m.ib[m.ib_cur];
	uint32_t* out = m.ib[m.ib_cur ^ 1];
	k_tri_keep<<<grid_for(tri_count), kThreads, 0, s>>>(in, tri_count, m.cremap, m.remap, m.tri_cnt);
	scan_exclusive(s, m.tri_cnt, m.tri_off, tri_count, nullptr, &m.sc->next_triangles, m.temp);
	k_tri_write<<<grid_for(tri_count), kThreads, 0, s>>>(in, tri_count, m.cremap, m.remap, m.tri_off, out);
	m.ib_cur ^= 1;
}

// ------------------------------------------------------------------------------------------
// finalisation (src/simplifier.cpp:2601-2619)

__global__ void k_debug_bits(uint32_t* ib, size_t tris, const uint8_t* kind, const uint32_t* loop, const uint32_t* loopback)
{
	size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= tris)
		return;
	uint32_t a = ib[t * 3 + 0], b = ib[t * 3 + 1], c = ib[t * 3 + 2];
	ib[t * 3 + 0] = a | ((uint32_t)kind[a] << 28) | ((uint32_t)(loop[a] == b || loopback[b] == a) << 31);
	ib[t * 3 + 1] = b | ((uint32_t)kind[b] << 28) | ((uint32_t)(loop[b] == c || loopback[c] == b) << 31);
	ib[t * 3 + 2] = c | ((uint32_t)kind[c] << 28) | ((uint32_t)(loop[c] == a || loopback[a] == c) << 31);
}

void launch_debug_bits(hipStream_t s, Mesh& m, size_t index_count)
{
	size_t tris = index_count / 3;
	if (tris)
		k_debug_bits<<<grid_for(tris), kThreads, 0, s>>>(m.ib[m.ib_cur], tris, m.kind, m.loop[m.loop_cur], m.loopback[m.loop_cur]);
}

__global__ void k_sparse_backmap(uint32_t* ib, size_t n, const uint32_t* sparse_remap)
{
	size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		ib[i] = sparse_remap[ib[i]];
}

void launch_sparse_backmap(hipStream_t s, uint32_t* ib, size_t index_count, const uint32_t* sparse_remap)
{
	if (index_count)
		k_sparse_backmap<<<grid_for(index_count), kThreads, 0, s>>>(ib, index_count, sparse_remap);
}

} // namespace mhip
