// Device code of the batched simplification of many small meshes (clusterlod-style patches):
// one workgroup per patch, the whole patch state resident in LDS, the entire simplify (setup,
// every pass, output) in one launch.  Each patch gets exactly what meshopt_simplify (src/simplifier.cpp:2628) returns
// for it: same indices, count and result_error.
//
// Parallel phases (remap, adjacency, classification, quadrics, pick, rank, sort, compaction) use
// the whole workgroup; the greedy commit (src/simplifier.cpp:1498-1636) is decided in rank order
// on wavefront 0, 64 candidates per step, exactly as the reference decides it.  All local ids
// are 16-bit.
#pragma once

#include "batch.h"
#include "quadric.h"

// threads per workgroup (a multiple of 64); a translation unit may build the device code with
// another count by defining MHIP_BATCH_THREADS first.  Everything below lives in an inline
// namespace named for the count, so two builds linked into one library stay apart.
#ifndef MHIP_BATCH_THREADS
#define MHIP_BATCH_THREADS 256
#endif
#define MHIP_BATCH_NS_CAT(a, b) a##b
#define MHIP_BATCH_NS(n) MHIP_BATCH_NS_CAT(threads, n)

namespace mhip
{
inline namespace MHIP_BATCH_NS(MHIP_BATCH_THREADS)
{
constexpr int kBatchThreads = MHIP_BATCH_THREADS;
static_assert(kBatchThreads % 64 == 0 && kBatchThreads >= 64 && kBatchThreads <= 1024, "workgroup of whole wavefronts");
constexpr int kBatchWaves = kBatchThreads / 64;

// kCanCollapse / kHasOpposite, src/simplifier.cpp:312-331, as 25-bit masks (bit k0 * 5 + k1)
// so that the lookups compile to shifts instead of per-lane memory loads
constexpr uint8_t kCanCollapseTable[5][5] = {
    {1, 1, 1, 1, 1},
    {0, 1, 0, 0, 1},
    {0, 0, 1, 0, 1},
    {0, 0, 0, 1, 1},
    {0, 0, 0, 0, 0},
};

constexpr uint8_t kHasOppositeTable[5][5] = {
    {1, 1, 1, 1, 1},
    {1, 0, 1, 0, 0},
    {1, 1, 1, 0, 1},
    {1, 0, 0, 0, 0},
    {1, 0, 1, 0, 0},
};

constexpr uint32_t kind_pair_mask(const uint8_t (&t)[5][5])
{
	uint32_t m = 0;
	for (int a = 0; a < 5; ++a)
		for (int b = 0; b < 5; ++b)
			m |= uint32_t(t[a][b] != 0) << (a * 5 + b);
	return m;
}

constexpr uint32_t kCanCollapseMask = kind_pair_mask(kCanCollapseTable);
constexpr uint32_t kHasOppositeMask = kind_pair_mask(kHasOppositeTable);

__device__ __forceinline__ uint32_t kind_pair(uint32_t mask, uint32_t k0, uint32_t k1)
{
	return (mask >> (k0 * 5 + k1)) & 1u;
}

constexpr uint16_t kNone16 = 0xffff;

// Optional phase timing (build with -DMESHOPT_BATCH_PROFILE): thread 0 of every workgroup adds
// the shader-clock cycles spent in each phase, and event counts, into g_batch_profile.
#ifdef MESHOPT_BATCH_PROFILE
__device__ unsigned long long g_batch_profile[32];
#define PROFILE_DECL unsigned long long prof_t = clock64(), prof_acc[32] = {}
#define PROFILE_PHASE(k)                          \
	do                                            \
	{                                             \
		unsigned long long now_ = clock64();      \
		prof_acc[k] += now_ - prof_t;             \
		prof_t = now_;                            \
	} while (0)
#define PROFILE_COUNT(k, n) (prof_acc[k] += (n))
#define PROFILE_FLUSH()                                        \
	do                                                         \
	{                                                          \
		if (threadIdx.x == 0)                                  \
			for (int k_ = 0; k_ < 32; ++k_)                    \
				atomicAdd(&g_batch_profile[k_], prof_acc[k_]); \
	} while (0)
#else
#define PROFILE_DECL
#define PROFILE_PHASE(k)
#define PROFILE_COUNT(k, n)
#define PROFILE_FLUSH()
#endif

// LDS layout of one patch, in bytes from the start of dynamic shared memory
struct PatchLayout
{
	uint32_t vmax, imax, cmax;
	uint32_t ib, px, py, pz, remap, wedge, loop, loopback, cremap, kind, lock, off, adj, sortbuf, region, cv0, cv1, order, cerr, total;
};

inline uint32_t align16(uint32_t x)
{
	return (x + 15) & ~15u;
}

// candidate arrays of a layout for meshes of up to imax indices: a pass has about one candidate
// per edge (I / 2 on a closed manifold, up to 0.58 I on open grids); a mesh with more leaves the
// kernel with status 5 and runs again with arrays of imax + 4, the bound boundEdgeCollapses gives
inline uint32_t batch_candidate_cap(uint32_t imax)
{
	return imax * 5 / 8 + 8;
}

inline PatchLayout patch_layout(uint32_t vmax, uint32_t imax, uint32_t cmax)
{
	PatchLayout L{};
	L.vmax = vmax;
	L.imax = imax;
	L.cmax = cmax;
	uint32_t o = 0;
	auto take = [&](uint32_t bytes) {
		uint32_t at = o;
		o = align16(o + bytes);
		return at;
	};
	L.ib = take(imax * 2);
	L.px = take(vmax * 4);
	L.py = take(vmax * 4);
	L.pz = take(vmax * 4);
	L.remap = take(vmax * 2);
	L.wedge = take(vmax * 2);
	L.loop = take(vmax * 2);
	L.loopback = take(vmax * 2);
	L.cremap = take(vmax * 2);
	L.kind = take(vmax);
	L.lock = take(vmax);
	L.off = take((vmax + 2) * 2);
	L.adj = take(imax * 4);
	// the sort's ping-pong half and digit counts (64 offsets + 64 counts per wavefront)
	L.sortbuf = take(align16(cmax * 2) + (1 + kBatchWaves) * 64 * 4);
	// one region for uses whose lifetimes do not overlap: the candidate arrays (pick to commit)
	// and the scratch of setup (position hash, wedge maxima), of build_adjacency (cursor, corner
	// ids; the ordered welded corners live on through the quadrics) and of the edge-loop update
	uint32_t hsize = 16;
	while (hsize < 2 * vmax + 2)
		hsize <<= 1;
	const uint32_t cand = 3 * align16(cmax * 2) + align16(cmax * 4);
	const uint32_t uses[4] = {4 * hsize, (vmax + 2) * 4 + imax * 2, 2 * align16(vmax * 2), vmax * 4};
	uint32_t rb = cand;
	for (uint32_t u : uses)
		rb = rb > u ? rb : u;
	L.region = take(rb);
	L.cv0 = L.region;
	L.cv1 = L.cv0 + align16(cmax * 2);
	L.order = L.cv1 + align16(cmax * 2);
	L.cerr = L.order + align16(cmax * 2);
	L.total = o;
	return L;
}

struct PatchSmem
{
	uint16_t *ib, *remap, *wedge, *loop, *loopback, *cremap, *off, *cv0, *cv1, *order, *sortbuf;
	float *px, *py, *pz, *cerr;
	Quad* gq; // the mesh's vertex quadrics in global memory (read and written by this workgroup only)
	uint8_t *kind, *lock;
	uint32_t* adj; // next | prev << 16
	uint8_t* scratch; // the region the candidate arrays share with setup and adjacency scratch
	uint32_t vmax, imax, cmax;
};

__device__ __forceinline__ PatchSmem bind(uint8_t* base, const PatchLayout& L)
{
	PatchSmem s;
	s.ib = (uint16_t*)(base + L.ib);
	s.px = (float*)(base + L.px);
	s.py = (float*)(base + L.py);
	s.pz = (float*)(base + L.pz);
	s.remap = (uint16_t*)(base + L.remap);
	s.wedge = (uint16_t*)(base + L.wedge);
	s.loop = (uint16_t*)(base + L.loop);
	s.loopback = (uint16_t*)(base + L.loopback);
	s.cremap = (uint16_t*)(base + L.cremap);
	s.kind = base + L.kind;
	s.lock = base + L.lock;
	s.off = (uint16_t*)(base + L.off);
	s.adj = (uint32_t*)(base + L.adj);
	s.sortbuf = (uint16_t*)(base + L.sortbuf);
	s.cv0 = (uint16_t*)(base + L.cv0);
	s.cv1 = (uint16_t*)(base + L.cv1);
	s.order = (uint16_t*)(base + L.order);
	s.cerr = (float*)(base + L.cerr);
	s.scratch = base + L.region;
	s.gq = nullptr;
	s.vmax = L.vmax;
	s.imax = L.imax;
	s.cmax = L.cmax;
	return s;
}

__device__ __forceinline__ V3 P(const PatchSmem& s, uint32_t i)
{
	return V3{s.px[i], s.py[i], s.pz[i]};
}

__device__ __forceinline__ Quad QL(const PatchSmem& s, uint32_t v)
{
	return load_quad(s.gq, v);
}

__device__ __forceinline__ void QS(const PatchSmem& s, uint32_t v, const Quad& q)
{
	store_quad(s.gq, v, q);
}

// exclusive scan of one value per thread across the workgroup; `part` holds kBatchThreads/64 words
__device__ __forceinline__ uint32_t wg_scan(uint32_t v, uint32_t* part, uint32_t& total)
{
	const int lane = __lane_id();
	uint32_t x = v;
#pragma unroll
	for (int o = 1; o < 64; o <<= 1)
	{
		uint32_t t = __shfl_up(x, o, 64);
		if (lane >= o)
			x += t;
	}
	const int wave = threadIdx.x >> 6;
	__syncthreads();
	if (lane == 63)
		part[wave] = x;
	__syncthreads();
	uint32_t before = 0, all = 0;
#pragma unroll
	for (int w = 0; w < kBatchThreads / 64; ++w)
	{
		uint32_t c = part[w];
		before += w < wave ? c : 0;
		all += c;
	}
	total = all;
	return before + x - v;
}

// counting placement of `count` items into per-key segments: off[key] = segment start, off[n] = total.
// keys come from key_of(i); with `ordered` the ids of the items in each segment come out in
// ascending item order, otherwise in any order.  cursor: n + 1 words of scratch.  All threads
// call it.
template <typename KeyOf>
__device__ void segment_items(uint16_t* off, uint16_t* items, uint32_t n, uint32_t count, KeyOf key_of, uint32_t* cursor, uint32_t* part,
                              bool ordered)
{
	for (uint32_t v = threadIdx.x; v <= n; v += kBatchThreads)
		cursor[v] = 0;
	__syncthreads();
	for (uint32_t i = threadIdx.x; i < count; i += kBatchThreads)
		atomicAdd(&cursor[key_of(i)], 1u);
	__syncthreads();
	// exclusive scan of cursor[0..n) (n <= a few thousand): chunks of kBatchThreads
	uint32_t carry = 0;
	for (uint32_t base = 0; base < n; base += kBatchThreads)
	{
		uint32_t v = base + threadIdx.x;
		uint32_t c = v < n ? cursor[v] : 0;
		uint32_t tot;
		uint32_t ex = wg_scan(c, part, tot);
		if (v < n)
			off[v] = (uint16_t)(carry + ex);
		carry += tot;
		__syncthreads();
	}
	if (threadIdx.x == 0)
		off[n] = (uint16_t)carry;
	__syncthreads();
	for (uint32_t v = threadIdx.x; v < n; v += kBatchThreads)
		cursor[v] = off[v];
	__syncthreads();
	for (uint32_t i = threadIdx.x; i < count; i += kBatchThreads)
		items[atomicAdd(&cursor[key_of(i)], 1u)] = (uint16_t)i;
	__syncthreads();
	if (!ordered)
		return;
	// ascending item order inside each segment (segments are short)
	for (uint32_t v = threadIdx.x; v < n; v += kBatchThreads)
	{
		uint32_t b = off[v], e = off[v + 1];
		for (uint32_t x = b + 1; x < e; ++x)
		{
			uint16_t t = items[x];
			uint32_t y = x;
			while (y > b && items[y - 1] > t)
			{
				items[y] = items[y - 1];
				--y;
			}
			items[y] = t;
		}
	}
	__syncthreads();
}

__device__ __forceinline__ bool lds_has_edge(const PatchSmem& s, uint32_t a, uint32_t b)
{
	for (uint32_t p = s.off[a], e = s.off[a + 1]; p < e; ++p)
		if ((s.adj[p] & 0xffffu) == b)
			return true;
	return false;
}

__device__ __forceinline__ bool lds_has_edge_wedged(const PatchSmem& s, uint32_t a, uint32_t b)
{
	uint32_t rb = s.remap[b];
	uint32_t v = a;
	do
	{
		for (uint32_t p = s.off[v], e = s.off[v + 1]; p < e; ++p)
			if (s.remap[s.adj[p] & 0xffffu] == rb)
				return true;
		v = s.wedge[v];
	} while (v != a);
	return false;
}

__device__ __forceinline__ uint32_t lds_complex_target(const PatchSmem& s, uint32_t v, uint32_t target)
{
	uint32_t r = s.remap[target];
	uint32_t l = s.loop[v];
	if (l != kNone16 && s.remap[l] == r)
		return l;
	uint32_t lb = s.loopback[v];
	if (lb != kNone16 && s.remap[lb] == r)
		return lb;
	return target;
}

__device__ __forceinline__ uint32_t canon_bits(float f)
{
	uint32_t u = __float_as_uint(f);
	return u == 0x80000000u ? 0u : u;
}

// build the corner CSR of the current index buffer keyed by `welded ? remap[v] : v`, each
// vertex's corners in corner order when `ordered` (the quadric sums need it; classification and
// the flip test only count, pick the single open edge, or OR over the corners);
// adj[p] = next | prev << 16 (both keyed the same way)
__device__ void build_adjacency(const PatchSmem& s, const uint16_t* ib, uint32_t icount, uint32_t vcount, bool welded, bool ordered,
                                uint16_t* corners, uint32_t* cursor, uint32_t* part)
{
	auto key_of = [&](uint32_t j) -> uint32_t {
		uint32_t v = ib[j];
		return welded ? s.remap[v] : v;
	};
	segment_items(s.off, corners, vcount, icount, key_of, cursor, part, ordered);
	for (uint32_t p = threadIdx.x; p < icount; p += kBatchThreads)
	{
		uint32_t j = corners[p];
		uint32_t t = j / 3, c = j - t * 3;
		uint32_t nx = ib[t * 3 + (c == 2 ? 0 : c + 1)];
		uint32_t pv = ib[t * 3 + (c == 0 ? 2 : c - 1)];
		if (welded)
		{
			nx = s.remap[nx];
			pv = s.remap[pv];
		}
		s.adj[p] = nx | (pv << 16);
	}
	__syncthreads();
}

// one mesh's attribute state in global memory: rescaled values, and per wedge the attribute
// quadric and one gradient per attribute
struct PatchAttr
{
	float* va;
	Quad* aq;
	Grad* ag;
	uint32_t count;
};

// quadricError with attributes (src/simplifier.cpp:784-799): wedge v's attribute quadric and
// gradients at wedge t's position and attribute values
__device__ __forceinline__ float patch_attr_error(const PatchAttr& at, const PatchSmem& s, uint32_t v, uint32_t t)
{
	return quad_error_attr(load_quad(at.aq, v), at.ag + (size_t)v * at.count, (int)at.count, P(s, t), at.va + (size_t)t * at.count);
}

// the attribute terms rankEdgeCollapses adds to a candidate v0 -> v1 (src/simplifier.cpp:1391-1431):
// ei (v0 -> v1) and, for a two-way edge, ej (v1 -> v0), with the seam pair's or the complex
// wedges' terms after the candidate's own
__device__ __forceinline__ void patch_attr_rank(const PatchAttr& at, const PatchSmem& s, uint32_t v0, uint32_t v1, bool bidi, float& ei, float& ej)
{
	ei += patch_attr_error(at, s, v0, v1);
	if (bidi)
		ej += patch_attr_error(at, s, v1, v0);
	const uint8_t k0 = s.kind[v0];
	if (k0 == kSeam)
	{
		const uint32_t s0 = s.wedge[v0];
		uint32_t s1 = s.loop[v0] == v1 ? s.loopback[s0] : s.loop[s0];
		s1 = s1 != kNone16 ? s1 : s.wedge[v1];
		ei += patch_attr_error(at, s, s0, s1);
		if (bidi)
			ej += patch_attr_error(at, s, s1, s0);
	}
	else
	{
		if (k0 == kComplex)
			for (uint32_t v = s.wedge[v0]; v != v0; v = s.wedge[v])
				ei += patch_attr_error(at, s, v, lds_complex_target(s, v, v1));
		if (s.kind[v1] == kComplex && bidi)
			for (uint32_t v = s.wedge[v1]; v != v1; v = s.wedge[v])
				ej += patch_attr_error(at, s, v, lds_complex_target(s, v, v0));
	}
}

// smallest member of r's wedge ring above `after` (kNone if none)
__device__ __forceinline__ uint32_t lds_next_wedge_above(const PatchSmem& s, uint32_t r, uint32_t after)
{
	uint32_t best = kNone;
	uint32_t v = r;
	do
	{
		if (v > after && v < best)
			best = v;
		v = s.wedge[v];
	} while (v != r);
	return best;
}

// The whole simplify of one mesh by the calling workgroup; every return is workgroup-uniform.
// sh[0..4] and shf[] are scratch words of this function.  kAttr: the batch has attributes
// (a.attribute_count > 0), which only that build of the function reads.
template <bool kAttr = false>
__device__ __forceinline__ void simplify_patch(const BatchArgs& a, const PatchLayout& L, uint32_t patch, uint8_t* smem, uint32_t* part,
                                               uint32_t* sh, uint64_t* box, float* shf)
{
	const uint32_t ib_begin = a.index_offsets[patch], ib_end = a.index_offsets[patch + 1];
	const uint32_t vb_begin = a.vertex_offsets[patch], vb_end = a.vertex_offsets[patch + 1];
	const uint32_t I = ib_end - ib_begin;
	const uint32_t V = vb_end - vb_begin;
	const uint32_t target = a.target_index_counts[patch];
	const uint32_t options = a.options;
	const int tid = threadIdx.x;

	if (V > L.vmax || I > L.imax || V >= kNone16)
	{
		if (tid == 0)
			a.status[patch] = 1; // does not fit: the host runs it through the plain path
		return;
	}

	PatchSmem s = bind(smem, L);
	s.gq = a.quadrics + vb_begin;
	PatchAttr at{};
	if constexpr (kAttr)
	{
		at.count = a.attribute_count;
		at.va = a.attr_values + (size_t)vb_begin * at.count;
		at.aq = a.attr_quadrics + vb_begin;
		at.ag = a.attr_grads + (size_t)vb_begin * at.count;
	}
	uint32_t* cursor = (uint32_t*)s.scratch; // V + 1 words
	PROFILE_DECL;

	// ---------------------------------------------------------------- load
	// each thread issues a batch of global loads before it uses any of them, so the memory
	// latency is paid once per batch rather than once per element
	if (tid == 0)
		sh[0] = 0;
	__syncthreads();
	{
		const size_t stride = a.position_stride_floats;
		constexpr int kPosBatch = 2, kIdxBatch = 8;
		for (uint32_t i0 = tid; i0 < V; i0 += kPosBatch * kBatchThreads)
		{
			float x[kPosBatch], y[kPosBatch], z[kPosBatch];
#pragma unroll
			for (int u = 0; u < kPosBatch; ++u)
			{
				const uint32_t i = i0 + u * kBatchThreads;
				x[u] = y[u] = z[u] = 0.f;
				if (i < V)
				{
					const float* p = a.positions + (size_t)(vb_begin + i) * stride;
					x[u] = p[0];
					y[u] = p[1];
					z[u] = p[2];
				}
			}
#pragma unroll
			for (int u = 0; u < kPosBatch; ++u)
			{
				const uint32_t i = i0 + u * kBatchThreads;
				if (i < V)
				{
					s.px[i] = x[u];
					s.py[i] = y[u];
					s.pz[i] = z[u];
				}
			}
		}
		bool bad = false;
		for (uint32_t j0 = tid; j0 < I; j0 += kIdxBatch * kBatchThreads)
		{
			uint32_t v[kIdxBatch];
#pragma unroll
			for (int u = 0; u < kIdxBatch; ++u)
			{
				const uint32_t j = j0 + u * kBatchThreads;
				v[u] = j < I ? a.indices[ib_begin + j] : 0u;
			}
#pragma unroll
			for (int u = 0; u < kIdxBatch; ++u)
			{
				const uint32_t j = j0 + u * kBatchThreads;
				if (j < I)
				{
					bad |= v[u] >= V;
					s.ib[j] = (uint16_t)(v[u] < V ? v[u] : 0);
				}
			}
		}
		if (bad)
			atomicOr(&sh[0], 1u);
	}
	__syncthreads();
	if (sh[0])
	{
		if (tid == 0)
			a.status[patch] = 2; // index out of range
		return;
	}

	// ---------------------------------------------------------------- position remap (src/simplifier.cpp:201-240)
	{
		uint32_t* table = (uint32_t*)s.scratch;
		uint32_t hsize = 16;
		while (hsize < 2 * V + 2)
			hsize <<= 1;
		const uint32_t mask = hsize - 1;
		for (uint32_t h = tid; h < hsize; h += kBatchThreads)
			table[h] = kNone;
		__syncthreads();
		for (uint32_t i = tid; i < V; i += kBatchThreads)
		{
			float x = s.px[i], y = s.py[i], z = s.pz[i];
			if (x != x || y != y || z != z)
				continue;
			uint64_t h = (uint64_t)canon_bits(x) * 0x9E3779B97F4A7C15ull ^ (uint64_t)canon_bits(y) * 0xC2B2AE3D27D4EB4Full ^
			             (uint64_t)canon_bits(z) * 0x165667B19E3779F9ull;
			uint32_t slot = ((uint32_t)h ^ (uint32_t)(h >> 32)) & mask;
			for (uint32_t probe = 0; probe <= mask; ++probe)
			{
				uint32_t old = atomicCAS(&table[slot], kNone, i);
				if (old == kNone)
					break;
				if (s.px[old] == x && s.py[old] == y && s.pz[old] == z)
				{
					atomicMin(&table[slot], i);
					break;
				}
				slot = (slot + 1) & mask;
			}
		}
		__syncthreads();
		for (uint32_t i = tid; i < V; i += kBatchThreads)
		{
			float x = s.px[i], y = s.py[i], z = s.pz[i];
			uint32_t r = i;
			if (!(x != x || y != y || z != z))
			{
				uint64_t h = (uint64_t)canon_bits(x) * 0x9E3779B97F4A7C15ull ^ (uint64_t)canon_bits(y) * 0xC2B2AE3D27D4EB4Full ^
				             (uint64_t)canon_bits(z) * 0x165667B19E3779F9ull;
				uint32_t slot = ((uint32_t)h ^ (uint32_t)(h >> 32)) & mask;
				for (uint32_t probe = 0; probe <= mask; ++probe)
				{
					uint32_t e = table[slot];
					if (e == kNone)
						break;
					if (s.px[e] == x && s.py[e] == y && s.pz[e] == z)
					{
						r = e;
						break;
					}
					slot = (slot + 1) & mask;
				}
			}
			s.remap[i] = (uint16_t)r;
		}
		__syncthreads();
		// wedge rings r -> a_m -> ... -> a_1 -> r: a secondary links to the next lower member
		uint32_t* wmax = (uint32_t*)s.scratch;
		for (uint32_t i = tid; i < V; i += kBatchThreads)
			wmax[i] = 0;
		__syncthreads();
		for (uint32_t i = tid; i < V; i += kBatchThreads)
		{
			uint32_t r = s.remap[i];
			uint32_t w = i;
			if (r != i)
			{
				uint32_t j = i - 1;
				while (s.remap[j] != r)
					--j;
				w = j;
				atomicMax(&wmax[r], i);
			}
			s.wedge[i] = (uint16_t)w;
		}
		__syncthreads();
		for (uint32_t i = tid; i < V; i += kBatchThreads)
			if (s.remap[i] == i && wmax[i])
				s.wedge[i] = (uint16_t)wmax[i];
		__syncthreads();
	}

	PROFILE_PHASE(0);
	// ---------------------------------------------------------------- unwelded adjacency, classification
	uint16_t* corners = (uint16_t*)(s.scratch + (V + 2) * 4);
	build_adjacency(s, s.ib, I, V, false, false, corners, cursor, part);

	for (uint32_t v = tid; v < V; v += kBatchThreads)
	{
		uint32_t out_count = 0, out_last = kNone, in_count = 0, in_last = kNone;
		bool self = false;
		for (uint32_t p = s.off[v], e = s.off[v + 1]; p < e; ++p)
		{
			uint32_t nx = s.adj[p] & 0xffffu, pv = s.adj[p] >> 16;
			if (nx == v)
				self = true;
			else if (!lds_has_edge(s, nx, v))
			{
				out_count++;
				out_last = nx;
			}
			if (pv == v)
				self = true;
			else if (!lds_has_edge(s, v, pv))
			{
				in_count++;
				in_last = pv;
			}
		}
		s.loop[v] = (uint16_t)((self || out_count > 1) ? v : (out_count == 1 ? out_last : kNone16));
		s.loopback[v] = (uint16_t)((self || in_count > 1) ? v : (in_count == 1 ? in_last : kNone16));
	}
	__syncthreads();

	for (uint32_t i = tid; i < V; i += kBatchThreads)
	{
		if (s.remap[i] != i)
			continue;
		uint8_t k;
		uint32_t oi = s.loopback[i], oo = s.loop[i];
		if (s.wedge[i] == i)
		{
			if (oi == kNone16 && oo == kNone16)
				k = kManifold;
			else if (oi != kNone16 && oo != kNone16 && s.remap[oi] == s.remap[oo] && oi != i)
				k = kSeam;
			else if (oi != i && oo != i)
				k = kBorder;
			else
				k = kLocked;
		}
		else if (s.wedge[s.wedge[i]] == i)
		{
			uint32_t w = s.wedge[i];
			uint32_t iw = s.loopback[w], ow = s.loop[w];
			bool open = oi != kNone16 && oi != i && oo != kNone16 && oo != i && iw != kNone16 && iw != w && ow != kNone16 && ow != w;
			k = (open && s.remap[oi] == s.remap[ow] && s.remap[oo] == s.remap[iw] && s.remap[oi] != s.remap[oo]) ? kSeam : kLocked;
		}
		else
			k = kLocked;
		s.kind[i] = k;
	}
	__syncthreads();
	if (options & kOptPermissive)
	{
		for (uint32_t i = tid; i < V; i += kBatchThreads)
		{
			if (s.remap[i] != i || (s.kind[i] != kSeam && s.kind[i] != kLocked))
				continue;
			bool protect = false;
			uint32_t v = i;
			// a Protect bit on any wedge keeps the vertex (src/simplifier.cpp:498-505)
			if (a.vertex_lock)
				do
				{
					protect |= (a.vertex_lock[vb_begin + v] & kVertexProtect) != 0;
					v = s.wedge[v];
				} while (v != i);
			do
			{
				for (uint32_t p = s.off[v], e = s.off[v + 1]; p < e && !protect; ++p)
					protect |= !lds_has_edge_wedged(s, s.adj[p] & 0xffffu, v);
				v = s.wedge[v];
			} while (v != i);
			if (!protect)
				s.kind[i] = kComplex;
		}
		__syncthreads();
	}
	// a Lock bit on any wedge locks its primary (src/simplifier.cpp:518-530); the loop below
	// carries the primary's kind to every wedge
	if (a.vertex_lock)
	{
		for (uint32_t i = tid; i < V; i += kBatchThreads)
			if (a.vertex_lock[vb_begin + i] & kVertexLock)
				s.kind[s.remap[i]] = kLocked;
		__syncthreads();
	}
	for (uint32_t i = tid; i < V; i += kBatchThreads)
	{
		uint8_t k = s.kind[s.remap[i]];
		if ((options & kOptLockBorder) && k == kBorder)
			k = kLocked;
		s.kind[i] = k;
	}
	__syncthreads();

	// boundEdgeCollapses (src/simplifier.cpp:1292-1308) on the unwelded adjacency
	uint32_t capacity;
	{
		uint32_t d = 0;
		for (uint32_t i = tid; i < V; i += kBatchThreads)
			d += (s.kind[i] == kManifold || s.kind[i] == kSeam) ? s.off[i + 1] - s.off[i] : 0u;
		uint32_t tot;
		wg_scan(d, part, tot);
		capacity = (I - tot / 2) + 3;
	}

	PROFILE_PHASE(1);
	// ---------------------------------------------------------------- rescale (src/simplifier.cpp:549-603)
	if (tid < 6)
		box[tid] = tid < 3 ? ~0ull : 0ull;
	__syncthreads();
	{
		uint64_t mn[3] = {~0ull, ~0ull, ~0ull}, mx[3] = {0, 0, 0};
		for (uint32_t i = tid; i < V; i += kBatchThreads)
		{
			float c[3] = {s.px[i], s.py[i], s.pz[i]};
			for (int k = 0; k < 3; ++k)
			{
				if (c[k] != c[k])
					continue;
				uint64_t key = (uint64_t)ordered_key(c[k]) << 32;
				uint64_t lo = key | i, hi = key | (0xffffffffu - i);
				mn[k] = lo < mn[k] ? lo : mn[k];
				mx[k] = hi > mx[k] ? hi : mx[k];
			}
		}
		// one LDS atomic per wavefront and bound
		for (int k = 0; k < 3; ++k)
		{
			const uint64_t wmn = wave_min64(mn[k]), wmx = wave_max64(mx[k]);
			if (__lane_id() == 0)
			{
				if (wmn != ~0ull)
					atomicMin((unsigned long long*)&box[k], (unsigned long long)wmn);
				if (wmx != 0ull)
					atomicMax((unsigned long long*)&box[3 + k], (unsigned long long)wmx);
			}
		}
	}
	__syncthreads();
	if (tid == 0)
	{
		float lo[3], hi[3];
		const float* comp[3] = {s.px, s.py, s.pz};
		for (int k = 0; k < 3; ++k)
		{
			lo[k] = FLT_MAX;
			hi[k] = -FLT_MAX;
			if (box[k] != ~0ull)
			{
				float v = comp[k][(uint32_t)box[k]];
				lo[k] = FLT_MAX > v ? v : FLT_MAX;
			}
			if (box[3 + k] != 0ull)
			{
				float v = comp[k][0xffffffffu - (uint32_t)box[3 + k]];
				hi[k] = -FLT_MAX < v ? v : -FLT_MAX;
			}
		}
		float extent = 0.f;
		extent = (hi[0] - lo[0]) < extent ? extent : (hi[0] - lo[0]);
		extent = (hi[1] - lo[1]) < extent ? extent : (hi[1] - lo[1]);
		extent = (hi[2] - lo[2]) < extent ? extent : (hi[2] - lo[2]);
		shf[0] = lo[0];
		shf[1] = lo[1];
		shf[2] = lo[2];
		shf[3] = extent == 0 ? 0.f : 1.f / extent;
		shf[4] = extent;
	}
	__syncthreads();
	for (uint32_t i = tid; i < V; i += kBatchThreads)
	{
		float sc = shf[3];
		s.px[i] = (s.px[i] - shf[0]) * sc;
		s.py[i] = (s.py[i] - shf[1]) * sc;
		s.pz[i] = (s.pz[i] - shf[2]) * sc;
	}
	const float extent = shf[4];
	__syncthreads();
	PROFILE_PHASE(12);

	if constexpr (kAttr)
	{
		// rescaleAttributes (src/simplifier.cpp:605-621): the attributes with weight > 0, each
		// times its weight
		const uint32_t A = at.count;
		for (uint32_t i = tid; i < V; i += kBatchThreads)
		{
			const float* src = a.attributes + (size_t)(vb_begin + i) * a.attribute_stride_floats;
			for (uint32_t k = 0; k < A; ++k)
				at.va[(size_t)i * A + k] = src[a.attr_map.index[k]] * a.attr_map.weight[k];
		}
		// fillAttributeQuadrics (src/simplifier.cpp:1159-1179): each wedge sums its triangles in
		// triangle order (the ordered corners of the unwelded adjacency; its barriers also make
		// the values above visible); the quadric first, then one gradient sum per attribute, so
		// that a single accumulator lives in registers
		build_adjacency(s, s.ib, I, V, false, true, corners, cursor, part);
		for (uint32_t v = tid; v < V; v += kBatchThreads)
		{
			const uint32_t b = s.off[v], e = s.off[v + 1];
			Quad Q;
			quad_zero(Q);
			for (uint32_t p = b; p < e; ++p)
			{
				const uint32_t t = corners[p] / 3;
				const uint32_t i0 = s.ib[t * 3 + 0], i1 = s.ib[t * 3 + 1], i2 = s.ib[t * 3 + 2];
				const AttrFrame f = attr_frame(P(s, i0), P(s, i1), P(s, i2));
				Quad QA;
				quad_zero(QA);
				QA.w = f.w;
				for (uint32_t k = 0; k < A; ++k)
				{
					Grad g;
					attr_term(QA, g, f, at.va[(size_t)i0 * A + k], at.va[(size_t)i1 * A + k], at.va[(size_t)i2 * A + k]);
				}
				quad_add(Q, QA);
			}
			store_quad(at.aq, v, Q);
			for (uint32_t k = 0; k < A; ++k)
			{
				Grad G{0.f, 0.f, 0.f, 0.f};
				for (uint32_t p = b; p < e; ++p)
				{
					const uint32_t t = corners[p] / 3;
					const uint32_t i0 = s.ib[t * 3 + 0], i1 = s.ib[t * 3 + 1], i2 = s.ib[t * 3 + 2];
					const AttrFrame f = attr_frame(P(s, i0), P(s, i1), P(s, i2));
					grad_add(G, attr_weighted_gradient(f, at.va[(size_t)i0 * A + k], at.va[(size_t)i1 * A + k], at.va[(size_t)i2 * A + k]));
				}
				at.ag[(size_t)v * A + k] = G;
			}
		}
		// the corner lists are rebuilt below
		__syncthreads();
	}

	// ---------------------------------------------------------------- vertex quadrics (src/simplifier.cpp:1057-1157)
	build_adjacency(s, s.ib, I, V, true, true, corners, cursor, part);
	PROFILE_PHASE(13);
	{
		const float factor = (options & kOptRegularizeLight) ? 1e-2f : ((options & kOptRegularize) ? 1e-1f : 1e-7f);
		for (uint32_t r = tid; r < V; r += kBatchThreads)
		{
			Quad Q;
			quad_zero(Q);
			if (s.remap[r] == r)
			{
				uint32_t b = s.off[r], e = s.off[r + 1];
				for (uint32_t p = b; p < e; ++p)
				{
					uint32_t t = corners[p] / 3;
					Quad F;
					quad_from_triangle(F, P(s, s.ib[t * 3 + 0]), P(s, s.ib[t * 3 + 1]), P(s, s.ib[t * 3 + 2]), 1.f);
					quad_add(Q, F);
				}
				{
					// Priority raises the regularising weight of a primary (src/simplifier.cpp:1095-1098)
					const bool priority = a.vertex_lock && (a.vertex_lock[vb_begin + r] & kVertexPriority) != 0;
					V3 v = P(s, r);
					Quad Pq;
					quad_from_point(Pq, v.x, v.y, v.z, Q.w * (priority ? 1.0f : factor));
					quad_add(Q, Pq);
				}
				uint32_t last_t = kNone;
				for (uint32_t p = b; p < e; ++p)
				{
					uint32_t t = corners[p] / 3;
					if (t == last_t)
						continue;
					last_t = t;
					uint32_t tri[3] = {s.ib[t * 3 + 0], s.ib[t * 3 + 1], s.ib[t * 3 + 2]};
					for (int ed = 0; ed < 3; ++ed)
					{
						uint32_t i0 = tri[ed], i1 = tri[ed == 2 ? 0 : ed + 1], i2 = tri[ed == 0 ? 2 : ed - 1];
						bool at0 = s.remap[i0] == r, at1 = s.remap[i1] == r;
						if (!at0 && !at1)
							continue;
						uint8_t k0 = s.kind[i0], k1 = s.kind[i1];
						bool b0 = k0 == kBorder || k0 == kSeam, b1 = k1 == kBorder || k1 == kSeam;
						if (!b0 && !b1)
							continue;
						if (b0 && s.loop[i0] != i1)
							continue;
						if (b1 && s.loopback[i1] != i0)
							continue;
						float weight = (k0 == kBorder || k1 == kBorder) ? 10.f : 0.5f;
						Quad E;
						quad_edge_term(E, P(s, i0), P(s, i1), P(s, i2), weight);
						if (at0)
							quad_add(Q, E);
						if (at1)
							quad_add(Q, E);
					}
				}
			}
			QS(s, r, Q);
		}
	}
	__syncthreads();

	PROFILE_PHASE(2);
	const float error_scale = (options & kOptErrorAbsolute) ? extent : 1.f;
	const float error_limit = (a.target_error * a.target_error) / (error_scale * error_scale);

	// ---------------------------------------------------------------- pass loop (src/simplifier.cpp:2493-2541)
	uint16_t* ib = s.ib;
	uint32_t result_count = I;
	float result_error = 0.f;
	uint32_t passes = 0;
	// per-pass totals for the caller's statistics: live triangles at the pass start, candidates,
	// committed collapses (workgroup-uniform values)
	uint32_t sum_live = 0, sum_candidates = 0, sum_committed = 0;

	// every pass that commits removes at least one triangle, so I / 3 + 1 passes always suffice;
	// the bound is an exit every wavefront reaches even on inconsistent state
	const uint32_t pass_limit = I / 3 + 1;
	bool overrun = false, overflow = false;
	while (result_count > target)
	{
		if (passes >= pass_limit)
		{
			overrun = true;
			break;
		}
		const uint32_t T = result_count / 3;
		build_adjacency(s, ib, result_count, V, true, false, corners, cursor, part);
		PROFILE_PHASE(3);

		// pick (src/simplifier.cpp:1310-1372) and rank (src/simplifier.cpp:1374-1450) in one pass over
		// the triangles: each thread finds its triangle's candidates, fetches the quadrics of the
		// corners that can be a collapse source while the edge checks run, and writes each taken
		// candidate once, oriented and with its error; order-preserving compaction
		uint32_t C = 0;
		if (tid == 0)
			sh[1] = 0;
		__syncthreads();
		{
			uint32_t carry = 0;
			for (uint32_t base = 0; base < T; base += kBatchThreads)
			{
				uint32_t t = base + tid;
				// up to three candidates per triangle, packed v0 | v1 << 16, in named registers
				uint32_t e0 = 0, e1 = 0, e2 = 0;
				float f0 = 0.f, f1 = 0.f, f2 = 0.f;
				int cnt = 0;
				if (t < T)
				{
					const uint32_t tri[3] = {ib[t * 3 + 0], ib[t * 3 + 1], ib[t * 3 + 2]};
					const uint32_t rv[3] = {s.remap[tri[0]], s.remap[tri[1]], s.remap[tri[2]]};
					const uint8_t kv[3] = {s.kind[tri[0]], s.kind[tri[1]], s.kind[tri[2]]};
					const V3 pv[3] = {P(s, tri[0]), P(s, tri[1]), P(s, tri[2])};
					// collapse errors of each corner toward its next and its previous corner (a
					// locked corner is never a collapse source); each quadric is dropped once used
					float to_next[3], to_prev[3];
#pragma unroll
					for (int k = 0; k < 3; ++k)
					{
						to_next[k] = to_prev[k] = 0.f;
						if (kv[k] != kLocked)
						{
							const Quad q = QL(s, rv[k]);
							to_next[k] = quad_error(q, pv[k == 2 ? 0 : k + 1]);
							to_prev[k] = quad_error(q, pv[k == 0 ? 2 : k - 1]);
						}
					}
#pragma unroll
					for (int ed = 0; ed < 3; ++ed)
					{
						const int a = ed, b = ed == 2 ? 0 : ed + 1;
						const uint32_t i0 = tri[a], i1 = tri[b];
						const uint32_t r0 = rv[a], r1 = rv[b];
						if (r0 == r1)
							continue;
						const uint8_t k0 = kv[a], k1 = kv[b];
						const uint32_t c01 = kind_pair(kCanCollapseMask, k0, k1), c10 = kind_pair(kCanCollapseMask, k1, k0);
						if (!(c01 | c10))
							continue;
						if (kind_pair(kHasOppositeMask, k0, k1) && r1 > r0)
							continue;
						if ((k0 == kBorder || k0 == kSeam) && k1 != kManifold && s.loop[i0] != i1)
							continue;
						if ((k1 == kBorder || k1 == kSeam) && k0 != kManifold && s.loopback[i1] != i0)
							continue;
						// the candidate as pickEdgeCollapses lists it (i0 -> i1 when both ways are
						// allowed), then rankEdgeCollapses: the cheaper direction of a two-way edge
						const bool bidi = c01 & c10;
						const bool fwd = bidi || c01;
						const uint32_t v0 = fwd ? i0 : i1, v1 = fwd ? i1 : i0;
						float ei = fwd ? to_next[a] : to_prev[b];
						float ej = bidi ? to_prev[b] : FLT_MAX;
						if constexpr (kAttr)
							patch_attr_rank(at, s, v0, v1, bidi, ei, ej);
						const bool rev = bidi & (ej < ei);
						const uint32_t packed = rev ? (v1 | (v0 << 16)) : (v0 | (v1 << 16));
						const float err = ej < ei ? ej : ei;
						if (cnt == 0)
						{
							e0 = packed;
							f0 = err;
						}
						else if (cnt == 1)
						{
							e1 = packed;
							f1 = err;
						}
						else
						{
							e2 = packed;
							f2 = err;
						}
						cnt++;
					}
				}
				uint32_t tot;
				uint32_t ex = wg_scan((uint32_t)cnt, part, tot);
				uint32_t at = carry + ex;
				// triangles are taken while `written + 3 <= capacity` (src/simplifier.cpp:1318-1320);
				// the running offset only grows, so the taken triangles form a prefix
				if (t < T && at + 3 <= capacity)
				{
					// past the layout's candidate arrays nothing is written; the count then
					// exceeds cmax and the mesh leaves with status 5
					if (cnt > 0 && at < s.cmax)
					{
						s.cv0[at] = (uint16_t)(e0 & 0xffffu);
						s.cv1[at] = (uint16_t)(e0 >> 16);
						s.cerr[at] = f0;
					}
					if (cnt > 1 && at + 1 < s.cmax)
					{
						s.cv0[at + 1] = (uint16_t)(e1 & 0xffffu);
						s.cv1[at + 1] = (uint16_t)(e1 >> 16);
						s.cerr[at + 1] = f1;
					}
					if (cnt > 2 && at + 2 < s.cmax)
					{
						s.cv0[at + 2] = (uint16_t)(e2 & 0xffffu);
						s.cv1[at + 2] = (uint16_t)(e2 >> 16);
						s.cerr[at + 2] = f2;
					}
					// the last taken triangle (the next one would not fit, or there is none) ends the list
					if (t + 1 == T || at + cnt + 3 > capacity)
						sh[1] = at + cnt;
				}
				carry += tot;
				__syncthreads();
			}
			C = sh[1];
			__syncthreads();
		}
		if (C > s.cmax)
		{
			overflow = true;
			break;
		}

		PROFILE_PHASE(4);
		PROFILE_COUNT(20, C);
		if (C == 0)
			break;

		PROFILE_PHASE(5);
		// stable sort by the 12-bit key (src/simplifier.cpp:1452-1496): LSD, two 6-bit digits
		{
			uint16_t* src = s.order;
			uint16_t* dst = s.sortbuf;
			for (uint32_t c = tid; c < C; c += kBatchThreads)
				src[c] = (uint16_t)c;
			__syncthreads();
			uint32_t* hist = (uint32_t*)((uint8_t*)s.sortbuf + ((s.cmax * 2 + 15) & ~15u)); // 64 digit offsets + [waves][64] counts
			for (int pass = 0; pass < 2; ++pass)
			{
				const int shift = pass * 6;
				for (uint32_t b = tid; b < 64; b += kBatchThreads)
					hist[b] = 0;
				__syncthreads();
				for (uint32_t c = tid; c < C; c += kBatchThreads)
					atomicAdd(&hist[(collapse_sort_key(s.cerr[src[c]]) >> shift) & 63u], 1u);
				__syncthreads();
				if (tid < 64)
				{
					uint32_t v = hist[tid];
					uint32_t x = v;
					for (int o = 1; o < 64; o <<= 1)
					{
						uint32_t t = __shfl_up(x, o, 64);
						if ((int)__lane_id() >= o)
							x += t;
					}
					hist[tid] = x - v;
				}
				__syncthreads();
				// stable scatter: chunks of kBatchThreads in order; rank among equal digits by ballots
				for (uint32_t base = 0; base < C; base += kBatchThreads)
				{
					uint32_t c = base + tid;
					bool valid = c < C;
					uint32_t item = valid ? src[c] : 0;
					uint32_t d = valid ? (collapse_sort_key(s.cerr[item]) >> shift) & 63u : 0;
					uint64_t peers = __ballot(valid);
					for (int b = 0; b < 6; ++b)
					{
						uint64_t bb = __ballot((d >> b) & 1u);
						peers &= ((d >> b) & 1u) ? bb : ~bb;
					}
					uint32_t lane = __lane_id();
					uint32_t rank_in_wave = (uint32_t)__popcll(peers & ((1ull << lane) - 1ull));
					uint32_t* wcount = hist + 64; // [kBatchWaves][64]
					const int wave = tid >> 6;
					for (uint32_t b = tid; b < kBatchWaves * 64; b += kBatchThreads)
						wcount[b] = 0;
					__syncthreads();
					if (valid && (63 - __clzll((long long)peers)) == (int)lane)
						wcount[wave * 64 + d] = (uint32_t)__popcll(peers);
					__syncthreads();
					uint32_t before = 0;
					for (int w = 0; w < wave; ++w)
						before += wcount[w * 64 + d];
					if (valid)
						dst[hist[d] + before + rank_in_wave] = (uint16_t)item;
					__syncthreads();
					if (tid < 64)
					{
						uint32_t all = 0;
						for (int w = 0; w < kBatchWaves; ++w)
							all += wcount[w * 64 + tid];
						hist[tid] += all;
					}
					__syncthreads();
				}
				uint16_t* t = src;
				src = dst;
				dst = t;
			}
			// after two passes the sorted order sits back in s.order
		}

		PROFILE_PHASE(6);
		// greedy commit (src/simplifier.cpp:1498-1636) on wavefront 0, 64 ranks at a time: every
		// lane loads one candidate and evaluates its lock state, its commit targets and its flip
		// test against the collapse_remap of the chunk start; the lanes are then resolved in rank
		// order with scalar state, exactly as the sequential loop decides them.  A lane's flip
		// result can only change if an earlier commit of the same chunk moved one of the welded
		// neighbours its test reads (the source class of that commit); each lane keeps its first
		// kNearSlots adjacency entries in registers to detect this exactly, and re-runs the test
		// on the current collapse_remap only then.
		for (uint32_t v = tid; v < V; v += kBatchThreads)
		{
			s.cremap[v] = (uint16_t)v;
			s.lock[v] = 0;
		}
		__syncthreads();
		if (tid < 64)
		{
			constexpr int kNearSlots = 8;
			const uint32_t lane = tid;
			const uint32_t goal = (result_count - target) / 3;
			uint32_t edge_goal = goal / 2;
			// cerr of the sorted candidates edge_goal .. edge_goal + 63, one per lane: the error
			// goal follows edge_goal one step per rejection and is read from this window
			uint32_t gbase = edge_goal;
			float gwin = gbase + lane < C ? s.cerr[s.order[gbase + lane]] : 0.f;
			float error_goal = edge_goal < C ? 1.5f * __int_as_float(__builtin_amdgcn_readlane(__float_as_int(gwin), 0)) : FLT_MAX;
			uint32_t tri_collapses = 0, committed = 0, evaluated = 0;
			bool stop = false;
			for (uint32_t base = 0; base < C && !stop; base += 64)
			{
				PROFILE_PHASE(15);
				PROFILE_COUNT(24, 1);
				const uint32_t k = base + lane;
				uint32_t i0 = 0, i1 = 0, r0 = kNone16, r1 = kNone16, s0 = kNone16, s1 = 0, ab = 0, ae = 0;
				float err = 0.f;
				bool locked = true, flips = false, border = false, complex = false;
				uint32_t near[kNearSlots];
#pragma unroll
				for (int q = 0; q < kNearSlots; ++q)
					near[q] = 0xffffffffu; // 0xffff never equals a vertex id
				if (k < C)
				{
					const uint32_t c = s.order[k];
					i0 = s.cv0[c];
					i1 = s.cv1[c];
					err = s.cerr[c];
					r0 = s.remap[i0];
					r1 = s.remap[i1];
					locked = (s.lock[r0] | s.lock[r1]) != 0;
					const uint8_t kd = s.kind[i0];
					border = kd == kBorder;
					complex = kd == kComplex;
					if (kd == kSeam)
					{
						s0 = s.wedge[i0];
						uint32_t t = s.loop[i0] == i1 ? s.loopback[s0] : s.loop[s0];
						s1 = t != kNone16 ? t : s.wedge[i1];
					}
					ab = s.off[r0];
					ae = s.off[r0 + 1];
					if (!locked)
					{
						// hasTriangleFlips (src/simplifier.cpp:1200-1232); a boolean OR over the
						// corners, evaluated without early exit so the loads overlap
						const V3 v0 = P(s, r0), v1 = P(s, r1);
#pragma unroll
						for (int q = 0; q < kNearSlots; ++q)
						{
							if (ab + q < ae)
							{
								const uint32_t ad = s.adj[ab + q];
								near[q] = ad;
								const uint32_t a2 = s.cremap[ad & 0xffffu], b2 = s.cremap[ad >> 16];
								if (!(a2 == r1 || b2 == r1 || a2 == b2))
									flips |= triangle_flips(P(s, a2), P(s, b2), v0, v1);
							}
						}
						for (uint32_t p = ab + kNearSlots; p < ae; ++p)
						{
							const uint32_t ad = s.adj[p];
							const uint32_t a2 = s.cremap[ad & 0xffffu], b2 = s.cremap[ad >> 16];
							if (!(a2 == r1 || b2 == r1 || a2 == b2))
								flips |= triangle_flips(P(s, a2), P(s, b2), v0, v1);
						}
					}
				}
				uint64_t lock_mask = __ballot(locked);
				const uint64_t flip_mask = __ballot(flips);
				// lanes whose corners do not all fit the register slots count as stale after any commit
				const uint64_t wide_mask = __ballot(ae - ab > (uint32_t)kNearSlots);
				uint64_t stale_mask = 0;
				const uint32_t n = C - base < 64 ? C - base : 64;
				const uint64_t valid = n == 64 ? ~0ull : (1ull << n) - 1;
				PROFILE_PHASE(14);
				// the stop test of every lane at the current scalar state; that state changes only on
				// a rejection or a commit, so the locked candidates between two such events are
				// passed over by one scan of the masks (they change nothing but the evaluated count)
				uint64_t stop_mask = __ballot(err > error_limit || tri_collapses >= goal || (err > error_goal && err > result_error && tri_collapses > goal / 6));
				for (uint32_t j = 0; j < n; ++j)
				{
					const uint64_t events = (stop_mask | ~lock_mask) & valid & (~0ull << j);
					if (!events)
					{
						evaluated += n - j;
						break;
					}
					const uint32_t jn = (uint32_t)__builtin_ctzll(events);
					evaluated += jn - j + 1;
					j = jn;
					if ((stop_mask >> j) & 1)
					{
						stop = true;
						break;
					}
					const float ej = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(err), j));
					const uint32_t r0j = __builtin_amdgcn_readlane(r0, j), r1j = __builtin_amdgcn_readlane(r1, j);
					bool fj = (flip_mask >> j) & 1;
					if ((stale_mask >> j) & 1)
					{
						// re-run on the current collapse_remap, one adjacency entry per lane
						PROFILE_COUNT(25, 1);
						PROFILE_PHASE(15);
						fj = false;
						const uint32_t b = __builtin_amdgcn_readlane(ab, j), e = __builtin_amdgcn_readlane(ae, j);
						const V3 v0 = P(s, r0j), v1 = P(s, r1j);
						for (uint32_t p0 = b; p0 < e; p0 += 64)
						{
							const uint32_t p = p0 + lane;
							bool f = false;
							if (p < e)
							{
								const uint32_t ad = s.adj[p];
								const uint32_t a2 = s.cremap[ad & 0xffffu], b2 = s.cremap[ad >> 16];
								if (!(a2 == r1j || b2 == r1j || a2 == b2))
									f = triangle_flips(P(s, a2), P(s, b2), v0, v1);
							}
							if (__ballot(f))
							{
								fj = true;
								break;
							}
						}
						PROFILE_PHASE(16);
					}
					if (fj)
					{
						edge_goal++;
						if (edge_goal - gbase >= 64)
						{
							gbase = edge_goal;
							gwin = gbase + lane < C ? s.cerr[s.order[gbase + lane]] : 0.f;
						}
						error_goal = edge_goal < C ? 1.5f * __int_as_float(__builtin_amdgcn_readlane(__float_as_int(gwin), edge_goal - gbase)) : FLT_MAX;
						stop_mask = __ballot(err > error_limit || tri_collapses >= goal || (err > error_goal && err > result_error && tri_collapses > goal / 6));
						continue;
					}
					// commit (src/simplifier.cpp:1573-1611) from the candidate's own lane
					if (lane == j)
					{
						if (complex)
						{
							uint32_t v = i0;
							do
							{
								s.cremap[v] = (uint16_t)lds_complex_target(s, v, i1);
								v = s.wedge[v];
							} while (v != i0);
						}
						else
						{
							s.cremap[i0] = (uint16_t)i1;
							if (s0 != kNone16)
								s.cremap[s0] = (uint16_t)s1;
						}
						s.lock[r0] = 1;
						s.lock[r1] = 1;
					}
					// the writes are read by later stale re-runs and by the next chunk's loads
					__builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
					__builtin_amdgcn_wave_barrier();
					lock_mask |= __ballot(r0 == r0j || r0 == r1j || r1 == r0j || r1 == r1j);
					bool moved = false;
#pragma unroll
					for (int q = 0; q < kNearSlots; ++q)
						moved |= (near[q] & 0xffffu) == r0j || (near[q] >> 16) == r0j;
					stale_mask |= __ballot(moved) | wide_mask;
					tri_collapses += __builtin_amdgcn_readlane((int)border, j) ? 1 : 2;
					committed++;
					result_error = result_error < ej ? ej : result_error;
					stop_mask = __ballot(err > error_limit || tri_collapses >= goal || (err > error_goal && err > result_error && tri_collapses > goal / 6));
				}
			}
			if (lane == 0)
			{
				sh[4] = evaluated;
				sh[3] = committed;
				shf[5] = result_error;
			}
		}
		__syncthreads();
		PROFILE_PHASE(7);
		PROFILE_COUNT(21, sh[4]);
		PROFILE_COUNT(22, sh[3]);
		result_error = shf[5];
		passes++;
		sum_live += T;
		sum_candidates += C;
		sum_committed += sh[3];
		if (sh[3] == 0)
			break;

		// updateQuadrics (src/simplifier.cpp:1638-1673): each target class receives one add; with
		// attributes, the moved wedges of the source class add their attribute quadrics and
		// gradients to their targets in ascending wedge order, as the reference's vertex loop
		// does (every wedge moving into a class comes from the one collapse that targets it)
		for (uint32_t i = tid; i < V; i += kBatchThreads)
		{
			uint32_t t = s.cremap[i];
			if (t != i && s.remap[i] == i)
			{
				Quad Q1 = QL(s, s.remap[t]);
				quad_add(Q1, QL(s, i));
				QS(s, s.remap[t], Q1);
				if constexpr (kAttr)
				{
					const uint32_t A = at.count;
					for (uint32_t v = i; v != kNone; v = lds_next_wedge_above(s, i, v))
					{
						const uint32_t tv = s.cremap[v];
						if (tv == v)
							continue;
						Quad Qt = load_quad(at.aq, tv);
						quad_add(Qt, load_quad(at.aq, v));
						store_quad(at.aq, tv, Qt);
						for (uint32_t k = 0; k < A; ++k)
						{
							Grad g = at.ag[(size_t)tv * A + k];
							grad_add(g, at.ag[(size_t)v * A + k]);
							at.ag[(size_t)tv * A + k] = g;
						}
					}
				}
			}
		}
		__syncthreads();

		// remapEdgeLoops (src/simplifier.cpp:1845-1864), double-buffered through scratch
		{
			uint16_t* nl = (uint16_t*)s.scratch;
			uint16_t* nb = nl + s.vmax;
			for (uint32_t i = tid; i < V; i += kBatchThreads)
			{
				const uint16_t* in[2] = {s.loop, s.loopback};
				uint16_t* out[2] = {nl, nb};
				for (int q = 0; q < 2; ++q)
				{
					uint32_t l = in[q][i];
					uint32_t res = kNone16;
					if (l != kNone16)
					{
						uint32_t r = s.cremap[l];
						if (i == r)
						{
							uint32_t ll = in[q][l];
							res = ll != kNone16 ? s.cremap[ll] : kNone16;
						}
						else
							res = r;
					}
					out[q][i] = (uint16_t)res;
				}
			}
			__syncthreads();
			for (uint32_t i = tid; i < V; i += kBatchThreads)
			{
				s.loop[i] = nl[i];
				s.loopback[i] = nb[i];
			}
			__syncthreads();
		}

		PROFILE_PHASE(8);
		// remapIndexBuffer (src/simplifier.cpp:1811-1843): stable compaction in place.  Each chunk's
		// triangles are read before the scan's barriers and written after them, at or below the
		// chunk's own range, so nothing unread is overwritten
		{
			uint32_t carry = 0;
			for (uint32_t base = 0; base < T; base += kBatchThreads)
			{
				uint32_t t = base + tid;
				uint32_t v[3] = {0, 0, 0};
				bool keep = false;
				if (t < T)
				{
					v[0] = s.cremap[ib[t * 3 + 0]];
					v[1] = s.cremap[ib[t * 3 + 1]];
					v[2] = s.cremap[ib[t * 3 + 2]];
					uint32_t q0 = s.remap[v[0]], q1 = s.remap[v[1]], q2 = s.remap[v[2]];
					keep = q0 != q1 && q0 != q2 && q1 != q2;
				}
				uint32_t tot;
				uint32_t ex = wg_scan(keep ? 1u : 0u, part, tot);
				if (keep)
				{
					uint32_t o = (carry + ex) * 3;
					ib[o + 0] = (uint16_t)v[0];
					ib[o + 1] = (uint16_t)v[1];
					ib[o + 2] = (uint16_t)v[2];
				}
				carry += tot;
				__syncthreads();
			}
			result_count = carry * 3;
		}
		__syncthreads();
		PROFILE_PHASE(9);
	}
	PROFILE_COUNT(23, passes);
	PROFILE_PHASE(10);

	if (overflow)
	{
		// more candidates than this layout holds: the host runs the mesh again with full arrays
		if (tid == 0)
			a.status[patch] = 5;
		return;
	}

	// ---------------------------------------------------------------- output (src/simplifier.cpp:2601-2625)
	const bool debug = (options & kOptInternalDebug) != 0;
	for (uint32_t j = tid; j < result_count; j += kBatchThreads)
	{
		uint32_t v = ib[j];
		if (debug)
		{
			uint32_t t = j / 3, c = j - t * 3;
			uint32_t nxt = ib[t * 3 + (c == 2 ? 0 : c + 1)];
			v |= ((uint32_t)s.kind[v] << 28) | ((uint32_t)(s.loop[v] == nxt || s.loopback[nxt] == v) << 31);
		}
		a.destination[ib_begin + j] = v;
	}
	if (tid == 0)
	{
		a.result_counts[patch] = result_count;
		if (a.result_errors)
			a.result_errors[patch] = sqrtf(result_error) * error_scale;
		a.status[patch] = overrun ? 3u : 0u;
		if (a.stats)
		{
			a.stats[patch * 4 + 0] = passes;
			a.stats[patch * 4 + 1] = sum_live;
			a.stats[patch * 4 + 2] = sum_candidates;
			a.stats[patch * 4 + 3] = sum_committed;
		}
	}
	PROFILE_PHASE(11);
	PROFILE_FLUSH();
}

// The meshes ids[0..n) taken from a cursor word shared by the launch (zero at its start): thread 0
// takes the next list position with one atomic and hands it to the workgroup through the LDS word
// `next`; the next mesh starts only after the barrier that ends the last one, which every thread
// reaches after reading `next`, so thread 0 never overwrites a value still unread.  The position
// read back is made wave-uniform (every thread reads the same word after the barrier, so the
// first lane's value is every lane's): the loop exit and every barrier loop inside simplify_patch
// then sit in uniform control flow, as in the one-mesh-per-workgroup form.
template <bool kAttr = false>
__device__ __forceinline__ void simplify_from_cursor(const BatchArgs& a, const PatchLayout& L, const uint32_t* ids, uint32_t n, uint32_t* cursor,
                                                     uint8_t* smem, uint32_t* part, uint32_t* sh, uint64_t* box, float* shf, uint32_t* next)
{
	for (;;)
	{
		if (threadIdx.x == 0)
			*next = atomicAdd(cursor, 1u);
		__syncthreads();
		const uint32_t i = __builtin_amdgcn_readfirstlane(*next);
		if (i >= n)
			break;
		simplify_patch<kAttr>(a, L, ids[i], smem, part, sh, box, shf);
		__syncthreads();
	}
}

} // inline namespace
} // namespace mhip
