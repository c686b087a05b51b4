// The vertex solve of meshopt_simplifyWithUpdate (src/simplifier.cpp:2575-2599) on the device.
//
// solvePositions (src/simplifier.cpp:1675-1763) walks the primaries in index order, and each
// vertex reads the current positions of its welded neighbours: a lower-index neighbour has
// already been moved (or kept), a higher-index one has not.  The walk is replayed exactly in
// rounds: a listed primary is solved in the first round in which every listed lower-index
// neighbour was solved in an earlier round.  Its higher-index neighbours wait for it, so
// nothing a round reads is written in the same round, and each vertex sees precisely the
// positions the sequential walk would show it.  Rounds = the longest index-increasing path
// through listed neighbours (tens for subdivided meshes, a few hundred for row-major grids).
//
// solveAttributes (1765-1809) and finalizeVertices (623-663) are independent per vertex.
#include "kernels.h"
#include "prims.h"
#include "quadric.h"

namespace mhip
{

constexpr uint32_t kPending = 0xffffffffu;
static constexpr int kThreads = 256;

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

__global__ void k_solve_flags(const uint8_t* update, const uint32_t* remap, const uint8_t* kind, size_t n, uint32_t* flags, uint32_t* state)
{
	size_t ii = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (ii >= n)
		return;
	uint32_t i = (uint32_t)ii;
	bool listed = update[i] && remap[i] == i && !solve_pinned(kind[i]);
	flags[i] = listed;
	state[i] = listed ? kPending : 0u;
}

__global__ void k_solve_list(const uint32_t* state, const uint32_t* offs, size_t n, uint32_t* list)
{
	size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n && state[i] == kPending)
		list[offs[i]] = (uint32_t)i;
}

__global__ void k_solve_reset(CallScalars* sc)
{
	sc->work[4] = 0;
	sc->work[5] = 0;
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
	k_solve_flags<<<grid_for(n), kThreads, 0, s>>>(b.update, m.remap, m.kind, n, b.flags, b.state);
	scan_exclusive(s, b.flags, b.flags, n, nullptr, &m.sc->work[4], m.temp);
	k_solve_list<<<grid_for(n), kThreads, 0, s>>>(b.state, b.flags, n, b.list);
}

// ------------------------------------------------------------------------------------------
// one round of solvePositions (src/simplifier.cpp:1675-1763)

struct SolveRound
{
	const uint32_t* list;
	uint32_t* state;
	float4* pos;
	const Quad* vq;
	const Grad* vg;
	const Quad* aq;
	const Grad* ag;
	int attr_count;
	const uint32_t* wedge;
	const uint32_t* adj_off;
	const uint2* adj;
	uint32_t* solved;
};

// the position the solve gives primary i, or false when one of the reference's checks keeps
// the old one (singular system, outside the neighbourhood, a flipped triangle, error growth)
__device__ bool solve_position(const SolveRound& a, uint32_t i, V3& out)
{
	const V3 cur = load_pos(a.pos, i);
	const Quad Qv = load_quad(a.vq, i);

	// a point quadric regularises the solution
	Quad Q = Qv;
	Quad R;
	quad_from_point(R, cur.x, cur.y, cur.z, Qv.w * 1e-4f);
	quad_add(Q, R);

	Grad GV = {0.f, 0.f, 0.f, 0.f};
	if (a.attr_count)
	{
		// the optimum minimises every wedge's attribute quadric at once
		uint32_t v = i;
		do
		{
			quad_reduce_attributes(Q, load_quad(a.aq, v), a.ag + (size_t)v * a.attr_count, a.attr_count);
			v = a.wedge[v];
		} while (v != i);
		// volume preservation as a constraint
		if (a.vg)
			GV = a.vg[i];
	}

	V3 p;
	if (!quad_solve(p, Q, GV))
		return false;

	const uint32_t begin = a.adj_off[i], end = a.adj_off[i + 1];

	// getNeighborhoodRadius, src/simplifier.cpp:1252-1276
	float far2 = 0.f;
	for (uint32_t e = begin; e < end; ++e)
	{
		uint2 ed = a.adj[e];
		V3 q0 = load_pos(a.pos, ed.x), q1 = load_pos(a.pos, ed.y);
		float d0 = (q0.x - cur.x) * (q0.x - cur.x) + (q0.y - cur.y) * (q0.y - cur.y) + (q0.z - cur.z) * (q0.z - cur.z);
		float d1 = (q1.x - cur.x) * (q1.x - cur.x) + (q1.y - cur.y) * (q1.y - cur.y) + (q1.z - cur.z) * (q1.z - cur.z);
		far2 = far2 < d0 ? d0 : far2;
		far2 = far2 < d1 ? d1 : far2;
	}
	float nr = sqrtf(far2);
	float dp = (p.x - cur.x) * (p.x - cur.x) + (p.y - cur.y) * (p.y - cur.y) + (p.z - cur.z) * (p.z - cur.z);
	if (dp > nr * nr)
		return false;

	// hasTriangleFlips for a point move, src/simplifier.cpp:1234-1250
	for (uint32_t e = begin; e < end; ++e)
	{
		uint2 ed = a.adj[e];
		if (triangle_flips(load_pos(a.pos, ed.x), load_pos(a.pos, ed.y), cur, p))
			return false;
	}

	if (quad_error(Qv, p) > quad_error(Qv, cur) * 1.5f + 1e-6f)
		return false;

	out = p;
	return true;
}

__global__ void __launch_bounds__(kThreads) k_solve_round(SolveRound a, uint32_t count, uint32_t round)
{
	const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
	uint32_t i = 0;
	bool ready = false;
	if (k < count)
	{
		i = a.list[k];
		ready = a.state[i] == kPending;
		// wait for every listed lower-index neighbour to be solved in an earlier round (pending
		// ones read all-ones, ones solved in this round read `round`)
		for (uint32_t e = a.adj_off[i], end = a.adj_off[i + 1]; ready && e < end; ++e)
		{
			uint2 ed = a.adj[e];
			if ((ed.x < i && a.state[ed.x] >= round) || (ed.y < i && a.state[ed.y] >= round))
				ready = false;
		}
	}
	if (ready)
	{
		V3 p;
		if (solve_position(a, i, p))
			a.pos[i] = make_float4(p.x, p.y, p.z, 0.f);
		a.state[i] = round;
	}
	const uint64_t done = __ballot(ready);
	if (__lane_id() == 0 && done)
		atomicAdd(a.solved, (uint32_t)__popcll(done));
}

void launch_solve_rounds(hipStream_t s, Mesh& m, const SolveBuffers& b, uint32_t list_count, uint32_t first, uint32_t rounds)
{
	if (!list_count)
		return;
	SolveRound a;
	a.list = b.list;
	a.state = b.state;
	a.pos = m.pos;
	a.vq = m.vq;
	a.vg = m.vg;
	a.aq = m.aq;
	a.ag = m.ag;
	a.attr_count = (int)m.attr_count;
	a.wedge = m.wedge;
	a.adj_off = m.adj_off;
	a.adj = m.adj;
	a.solved = &m.sc->work[5];
	for (uint32_t r = 0; r < rounds; ++r)
		k_solve_round<<<grid_for(list_count), kThreads, 0, s>>>(a, list_count, first + r);
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
	k_solve_wedges<<<grid_for(n), kThreads, 0, s>>>(b.update, m.remap, m.kind, n, m.pos);
	if (m.attr_count)
		k_solve_attributes<<<grid_for(n), kThreads, 0, s>>>(b.update, m.remap, m.wedge, m.kind, m.pos, m.aq, m.ag, (int)m.attr_count, n, m.va);
	k_solve_finalize<<<grid_for(n), kThreads, 0, s>>>(b.update, m.kind, m.pos, m.va, (int)m.attr_count, n, m.sc, o);
}

} // namespace mhip
