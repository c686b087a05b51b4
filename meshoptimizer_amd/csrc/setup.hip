// Setup kernels of the edge-collapse simplifier: sparse remap, position remap and wedge
// rings, corner CSR, vertex classification, rescale, and the ordered quadric gathers.
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

// number of bits needed to hold values < n
static int key_bits(size_t n)
{
	int b = 0;
	while (b < 32 && ((size_t)1 << b) < n)
		++b;
	return b;
}

// ------------------------------------------------------------------------------------------
// input validation: any index >= vertex_count sets the flag

__global__ void k_check_indices(const uint32_t* ib, size_t n, size_t limit, uint32_t* flag)
{
	size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	bool bad = j < n && ib[j] >= limit;
	if (__ballot(bad) && __lane_id() == 0)
		atomicOr(flag, 1u);
}

void launch_check_indices(hipStream_t s, const uint32_t* ib, size_t index_count, size_t vertex_count, uint32_t* flag)
{
	if (index_count)
		k_check_indices<<<grid_for(index_count), kThreads, 0, s>>>(ib, index_count, vertex_count, flag);
}

// ------------------------------------------------------------------------------------------
// sparse remap (src/simplifier.cpp:242-294): dense id = order of first occurrence

__global__ void k_first_occurrence(const uint32_t* ib, size_t n, uint32_t* first)
{
	size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (j < n)
		atomicMin(&first[ib[j]], (uint32_t)j);
}

__global__ void k_first_flags(const uint32_t* ib, size_t n, const uint32_t* first, uint32_t* flags)
{
	size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (j < n)
		flags[j] = first[ib[j]] == (uint32_t)j;
}

__global__ void k_sparse_write(const uint32_t* ib_in, uint32_t* ib_out, size_t n, const uint32_t* first, const uint32_t* dense, uint32_t* sparse_remap)
{
	size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (j >= n)
		return;
	uint32_t v = ib_in[j];
	uint32_t f = first[v];
	if (f == (uint32_t)j)
		sparse_remap[dense[j]] = v;
	ib_out[j] = dense[f];
}

void launch_sparse_remap(hipStream_t s, const uint32_t* indices_in, uint32_t* indices_out, size_t index_count, size_t vertex_count,
                         uint32_t* first, uint32_t* flags, uint32_t* dense, uint32_t* sparse_remap, uint32_t* unique_count, uint32_t* temp)
{
	fill_u32(s, first, kNone, vertex_count);
	if (index_count == 0)
	{
		HIP_CHECK(hipMemsetAsync(unique_count, 0, sizeof(uint32_t), s));
		return;
	}
	k_first_occurrence<<<grid_for(index_count), kThreads, 0, s>>>(indices_in, index_count, first);
	k_first_flags<<<grid_for(index_count), kThreads, 0, s>>>(indices_in, index_count, first, flags);
	scan_exclusive(s, flags, dense, index_count, nullptr, unique_count, temp);
	k_sparse_write<<<grid_for(index_count), kThreads, 0, s>>>(indices_in, indices_out, index_count, first, dense, sparse_remap);
}

// ------------------------------------------------------------------------------------------
// positions

__global__ void k_gather_positions(const float* positions, size_t stride, const uint32_t* sparse_remap, float4* out, size_t n)
{
	size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	size_t ri = sparse_remap ? sparse_remap[i] : i;
	const float* p = positions + ri * stride;
	out[i] = make_float4(p[0], p[1], p[2], 0.f);
}

void launch_gather_positions(hipStream_t s, const float* positions, size_t stride_floats, const uint32_t* sparse_remap, float4* out, size_t vertex_count)
{
	if (vertex_count)
		k_gather_positions<<<grid_for(vertex_count), kThreads, 0, s>>>(positions, stride_floats, sparse_remap, out, vertex_count);
}

// ------------------------------------------------------------------------------------------
// position remap (src/simplifier.cpp:111-240).  remap[i] = lowest index whose xyz compare
// equal as floats (+0 == -0; NaN equals nothing).  Linear-probing table keyed by the
// canonical bits; a slot holds one equality class, and atomicMin leaves its lowest index.

__device__ __forceinline__ bool has_nan(float4 p)
{
	return p.x != p.x || p.y != p.y || p.z != p.z;
}

__device__ __forceinline__ uint32_t canon(float f)
{
	uint32_t u = __float_as_uint(f);
	return u == 0x80000000u ? 0u : u;
}

__device__ __forceinline__ uint32_t pos_hash(float4 p)
{
	uint64_t h = (uint64_t)canon(p.x) * 0x9E3779B97F4A7C15ull;
	h ^= (uint64_t)canon(p.y) * 0xC2B2AE3D27D4EB4Full;
	h ^= (uint64_t)canon(p.z) * 0x165667B19E3779F9ull;
	h ^= h >> 29;
	return (uint32_t)h ^ (uint32_t)(h >> 32);
}

__device__ __forceinline__ bool same_pos(float4 a, float4 b)
{
	return a.x == b.x && a.y == b.y && a.z == b.z;
}

__global__ void k_hash_insert(const float4* pos, size_t n, uint32_t* table, uint32_t mask)
{
	size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	float4 p = pos[i];
	if (has_nan(p))
		return;
	uint32_t slot = pos_hash(p) & mask;
	for (uint32_t probe = 0; probe <= mask; ++probe)
	{
		uint32_t old = atomicCAS(&table[slot], kNone, (uint32_t)i);
		if (old == kNone)
			return;
		if (same_pos(pos[old], p))
		{
			atomicMin(&table[slot], (uint32_t)i);
			return;
		}
		slot = (slot + 1) & mask;
	}
}

__global__ void k_hash_lookup(const float4* pos, size_t n, const uint32_t* table, uint32_t mask, uint32_t* remap)
{
	size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	float4 p = pos[i];
	uint32_t result = (uint32_t)i;
	if (!has_nan(p))
	{
		uint32_t slot = pos_hash(p) & mask;
		for (uint32_t probe = 0; probe <= mask; ++probe)
		{
			uint32_t e = table[slot];
			if (e == kNone)
				break; // unreachable: i itself was inserted
			if (same_pos(pos[e], p))
			{
				result = e;
				break;
			}
			slot = (slot + 1) & mask;
		}
	}
	remap[i] = result;
}

void launch_position_remap(hipStream_t s, Mesh& m, uint32_t* table, size_t table_size)
{
	size_t n = m.vertex_count;
	if (!n)
		return;
	fill_u32(s, table, kNone, table_size);
	k_hash_insert<<<grid_for(n), kThreads, 0, s>>>(m.praw, n, table, (uint32_t)(table_size - 1));
	k_hash_lookup<<<grid_for(n), kThreads, 0, s>>>(m.praw, n, table, (uint32_t)(table_size - 1), m.remap);
}

// wedge rings: for a class r < a_1 < ... < a_m, wedge[r] = a_m, wedge[a_k] = a_{k-1}, wedge[a_1] = r
// (the list buildPositionRemap forms by inserting each secondary after its primary,
// src/simplifier.cpp:229-239).  Secondaries are compacted in index order, stably sorted by
// class, and linked to their predecessor.

__global__ void k_secondary_flags(const uint32_t* remap, size_t n, uint32_t* flags)
{
	size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		flags[i] = remap[i] != (uint32_t)i;
}

__global__ void k_secondary_compact(const uint32_t* remap, size_t n, const uint32_t* flags, const uint32_t* offs, uint32_t* keys, uint32_t* vals)
{
	size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n && flags[i])
	{
		keys[offs[i]] = remap[i];
		vals[offs[i]] = (uint32_t)i;
	}
}

__global__ void k_wedge_link(const uint32_t* keys, const uint32_t* vals, size_t n_max, const uint32_t* n_dev, uint32_t* wedge)
{
	size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	size_t n = *n_dev < n_max ? *n_dev : n_max;
	if (p >= n)
		return;
	uint32_t r = keys[p];
	uint32_t i = vals[p];
	wedge[i] = (p > 0 && keys[p - 1] == r) ? vals[p - 1] : r;
	if (p + 1 == n || keys[p + 1] != r)
		wedge[r] = i;
}

void launch_wedges(hipStream_t s, Mesh& m, uint32_t* flags, uint32_t* scratch_a, uint32_t* scratch_b, uint32_t* scratch_c, uint32_t* scratch_d)
{
	size_t n = m.vertex_count;
	if (!n)
		return;
	iota_u32(s, m.wedge, n);
	uint32_t* count = &m.sc->work[0];
	k_secondary_flags<<<grid_for(n), kThreads, 0, s>>>(m.remap, n, flags);
	scan_exclusive(s, flags, scratch_a, n, nullptr, count, m.temp);
	k_secondary_compact<<<grid_for(n), kThreads, 0, s>>>(m.remap, n, flags, scratch_a, scratch_b, scratch_c);
	radix_sort_pairs(s, scratch_b, scratch_c, scratch_d, scratch_a, n, count, key_bits(n), m.temp);
	k_wedge_link<<<grid_for(n), kThreads, 0, s>>>(scratch_b, scratch_c, n, count, m.wedge);
}

// ------------------------------------------------------------------------------------------
// corner CSR sorted by (vertex, corner): the order in which the reference visits a vertex's
// triangles (triangle order, then corner order)

__global__ void k_corner_keys(const uint32_t* ib, size_t n, const uint32_t* remap, uint32_t* keys, uint32_t* vals, uint32_t* counts)
{
	size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (j >= n)
		return;
	uint32_t v = ib[j];
	if (remap)
		v = remap[v];
	keys[j] = v;
	vals[j] = (uint32_t)j;
	atomicAdd(&counts[v], 1u);
}

void launch_corner_csr(hipStream_t s, const uint32_t* ib, size_t index_count, const uint32_t* remap, size_t vertex_count,
                       uint32_t* offsets, uint32_t* keys, uint32_t* corners, uint32_t* keys_alt, uint32_t* vals_alt, uint32_t* temp)
{
	fill_u32(s, offsets, 0, vertex_count + 1);
	if (index_count == 0)
		return;
	k_corner_keys<<<grid_for(index_count), kThreads, 0, s>>>(ib, index_count, remap, keys, corners, offsets);
	scan_exclusive(s, offsets, offsets, vertex_count, nullptr, offsets + vertex_count, temp);
	radix_sort_pairs(s, keys, corners, keys_alt, vals_alt, index_count, nullptr, key_bits(vertex_count), temp);
}

// ------------------------------------------------------------------------------------------
// unwelded half-edge data (next, prev) per sorted corner

__global__ void k_corner_adjacency(const uint32_t* ib, const uint32_t* corners, size_t n, uint2* adj)
{
	size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (p >= n)
		return;
	uint32_t j = corners[p];
	uint32_t t = j / 3, c = j - t * 3;
	uint32_t cn = c == 2 ? 0 : c + 1;
	uint32_t cp = c == 0 ? 2 : c - 1;
	adj[p] = make_uint2(ib[t * 3 + cn], ib[t * 3 + cp]);
}

// hasEdge, src/simplifier.cpp:333-343
__device__ __forceinline__ bool has_edge(const uint32_t* off, const uint2* adj, uint32_t a, uint32_t b)
{
	for (uint32_t p = off[a], e = off[a + 1]; p < e; ++p)
		if (adj[p].x == b)
			return true;
	return false;
}

// hasEdge over the wedges of a, src/simplifier.cpp:345-362
__device__ __forceinline__ bool has_edge_wedged(const uint32_t* off, const uint2* adj, uint32_t a, uint32_t b, const uint32_t* remap, const uint32_t* wedge)
{
	uint32_t rb = remap[b];
	uint32_t v = a;
	do
	{
		for (uint32_t p = off[v], e = off[v + 1]; p < e; ++p)
			if (remap[adj[p].x] == rb)
				return true;
		v = wedge[v];
	} while (v != a);
	return false;
}

// open half-edges per vertex (src/simplifier.cpp:379-400), gathered from the vertex's own
// corners: an outgoing edge v->t is open when t has no edge t->v, an incoming edge u->v is
// open when v has no edge v->u.  Result: none, the single other endpoint, or v itself when
// there are several open edges or a self edge.
__global__ void k_open_edges(const uint32_t* off, const uint2* adj, size_t n, uint32_t* openout, uint32_t* openinc)
{
	size_t iv = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (iv >= n)
		return;
	uint32_t v = (uint32_t)iv;
	uint32_t out_count = 0, out_last = kNone, in_count = 0, in_last = kNone;
	bool self = false;
	for (uint32_t p = off[v], e = off[v + 1]; p < e; ++p)
	{
		uint2 np = adj[p];
		if (np.x == v)
			self = true;
		else if (!has_edge(off, adj, np.x, v))
		{
			out_count++;
			out_last = np.x;
		}
		if (np.y == v)
			self = true;
		else if (!has_edge(off, adj, v, np.y))
		{
			in_count++;
			in_last = np.y;
		}
	}
	openout[v] = (self || out_count > 1) ? v : (out_count == 1 ? out_last : kNone);
	openinc[v] = (self || in_count > 1) ? v : (in_count == 1 ? in_last : kNone);
}

// classifyVertices primary rules, src/simplifier.cpp:406-470
__global__ void k_classify_primary(const uint32_t* remap, const uint32_t* wedge, const uint32_t* loop, const uint32_t* loopback, size_t n, uint8_t* kind)
{
	size_t ii = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (ii >= n)
		return;
	uint32_t i = (uint32_t)ii;
	if (remap[i] != i)
		return;

	uint8_t k;
	if (wedge[i] == i)
	{
		uint32_t oi = loopback[i], oo = loop[i];
		if (oi == kNone && oo == kNone)
			k = kManifold;
		else if (oi != kNone && oo != kNone && remap[oi] == remap[oo] && oi != i)
			k = kSeam;
		else if (oi != i && oo != i)
			k = kBorder;
		else
			k = kLocked;
	}
	else if (wedge[wedge[i]] == i)
	{
		uint32_t w = wedge[i];
		uint32_t iv = loopback[i], ov = loop[i];
		uint32_t iw = loopback[w], ow = loop[w];
		bool open = iv != kNone && iv != i && ov != kNone && ov != i && iw != kNone && iw != w && ow != kNone && ow != w;
		if (open && remap[iv] == remap[ow] && remap[ov] == remap[iw] && remap[iv] != remap[ov])
			k = kSeam;
		else
			k = kLocked;
	}
	else
		k = kLocked;
	kind[i] = k;
}

// permissive mode, src/simplifier.cpp:481-515 (primaries; wedges copy later)
__global__ void k_permissive(const uint32_t* off, const uint2* adj, const uint32_t* remap, const uint32_t* wedge, const uint8_t* vertex_lock,
                             const uint32_t* sparse_remap, size_t n, uint8_t* kind)
{
	size_t ii = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (ii >= n)
		return;
	uint32_t i = (uint32_t)ii;
	if (remap[i] != i || (kind[i] != kSeam && kind[i] != kLocked))
		return;

	bool protect = false;
	uint32_t v = i;
	do
	{
		uint32_t rv = sparse_remap ? sparse_remap[v] : v;
		protect |= vertex_lock && (vertex_lock[rv] & kVertexProtect) != 0;
		v = wedge[v];
	} while (v != i);

	do
	{
		for (uint32_t p = off[v], e = off[v + 1]; p < e && !protect; ++p)
			protect |= !has_edge_wedged(off, adj, adj[p].x, v, remap, wedge);
		v = wedge[v];
	} while (v != i);

	if (!protect)
		kind[i] = kComplex;
}

// vertex_lock Lock bit, src/simplifier.cpp:517-525 (marks the primary)
__global__ void k_lock_mark(const uint32_t* remap, const uint8_t* vertex_lock, const uint32_t* sparse_remap, size_t n, uint8_t* kind)
{
	size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	size_t ri = sparse_remap ? sparse_remap[i] : i;
	if (vertex_lock[ri] & kVertexLock)
		kind[remap[i]] = kLocked;
}

// every wedge takes its primary's final kind; LockBorder turns Border into Locked
// (src/simplifier.cpp:470-477, 527-536; the Border->Locked map is idempotent, so primaries
// updating in place race harmlessly with wedges reading them)
__global__ void k_kind_finalize(const uint32_t* remap, size_t n, uint32_t options, uint8_t* kind)
{
	size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	uint8_t k = kind[remap[i]];
	if ((options & kOptLockBorder) && k == kBorder)
		k = kLocked;
	kind[i] = k;
}

void launch_classify(hipStream_t s, Mesh& m, const uint32_t* raw_off, const uint32_t* raw_corners, const uint8_t* vertex_lock,
                     const uint32_t* sparse_remap, uint32_t options)
{
	size_t n = m.vertex_count;
	size_t ni = m.index_count;
	if (!n)
		return;
	uint32_t* loop = m.loop[0];
	uint32_t* loopback = m.loopback[0];
	if (ni)
		k_corner_adjacency<<<grid_for(ni), kThreads, 0, s>>>(m.ib[0], raw_corners, ni, m.adj);
	k_open_edges<<<grid_for(n), kThreads, 0, s>>>(raw_off, m.adj, n, loop, loopback);
	k_classify_primary<<<grid_for(n), kThreads, 0, s>>>(m.remap, m.wedge, loop, loopback, n, m.kind);
	if (options & kOptPermissive)
		k_permissive<<<grid_for(n), kThreads, 0, s>>>(raw_off, m.adj, m.remap, m.wedge, vertex_lock, sparse_remap, n, m.kind);
	if (vertex_lock)
		k_lock_mark<<<grid_for(n), kThreads, 0, s>>>(m.remap, vertex_lock, sparse_remap, n, m.kind);
	k_kind_finalize<<<grid_for(n), kThreads, 0, s>>>(m.remap, n, options, m.kind);
}

// ------------------------------------------------------------------------------------------
// rescale (src/simplifier.cpp:549-603).  The reference keeps the first element (in vertex
// order) that compares below the running minimum, so among +0/-0 ties the earliest wins:
// reduce (ordered key, index) pairs, then read the winning element's own bits.

__global__ void __launch_bounds__(256) k_bbox(const float4* praw, size_t n, CallScalars* sc)
{
	__shared__ uint64_t red[6][4];
	uint64_t mn[3], mx[3];
#pragma unroll
	for (int k = 0; k < 3; ++k)
	{
		mn[k] = ~0ull;
		mx[k] = 0ull;
	}
	for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
	{
		float4 p = praw[i];
		float c[3] = {p.x, p.y, p.z};
#pragma unroll
		for (int k = 0; k < 3; ++k)
		{
			if (c[k] != c[k])
				continue;
			uint64_t key = (uint64_t)ordered_key(c[k]) << 32;
			uint64_t a = key | (uint64_t)(uint32_t)i;
			uint64_t b = key | (uint64_t)(0xffffffffu - (uint32_t)i);
			mn[k] = a < mn[k] ? a : mn[k];
			mx[k] = b > mx[k] ? b : mx[k];
		}
	}
	const int wave = threadIdx.x >> 6;
#pragma unroll
	for (int k = 0; k < 3; ++k)
	{
		uint64_t a = wave_min64(mn[k]);
		uint64_t b = wave_max64(mx[k]);
		if (__lane_id() == 0)
		{
			red[k][wave] = a;
			red[3 + k][wave] = b;
		}
	}
	__syncthreads();
	if (threadIdx.x < 3)
	{
		int k = threadIdx.x;
		uint64_t a = ~0ull, b = 0ull;
		for (int w = 0; w < 4; ++w)
		{
			a = red[k][w] < a ? red[k][w] : a;
			b = red[3 + k][w] > b ? red[3 + k][w] : b;
		}
		if (a != ~0ull)
			atomicMin((unsigned long long*)&sc->box_min[k], (unsigned long long)a);
		if (b != 0ull)
			atomicMax((unsigned long long*)&sc->box_max[k], (unsigned long long)b);
	}
}

__global__ void k_bbox_init(CallScalars* sc)
{
	for (int k = 0; k < 3; ++k)
	{
		sc->box_min[k] = ~0ull;
		sc->box_max[k] = 0ull;
	}
}

__device__ __forceinline__ float comp(float4 p, int k)
{
	return k == 0 ? p.x : (k == 1 ? p.y : p.z);
}

__global__ void k_extent(const float4* praw, CallScalars* sc)
{
	float lo[3], hi[3];
	for (int k = 0; k < 3; ++k)
	{
		lo[k] = FLT_MAX;
		hi[k] = -FLT_MAX;
		if (sc->box_min[k] != ~0ull)
		{
			float v = comp(praw[(uint32_t)sc->box_min[k]], k);
			lo[k] = FLT_MAX > v ? v : FLT_MAX;
		}
		if (sc->box_max[k] != 0ull)
		{
			float v = comp(praw[0xffffffffu - (uint32_t)sc->box_max[k]], k);
			hi[k] = -FLT_MAX < v ? v : -FLT_MAX;
		}
	}

	float extent = 0.f;
	extent = (hi[0] - lo[0]) < extent ? extent : (hi[0] - lo[0]);
	extent = (hi[1] - lo[1]) < extent ? extent : (hi[1] - lo[1]);
	extent = (hi[2] - lo[2]) < extent ? extent : (hi[2] - lo[2]);

	for (int k = 0; k < 3; ++k)
		sc->lo[k] = lo[k];
	sc->extent = extent;
	sc->scale = extent == 0 ? 0.f : 1.f / extent;
}

__global__ void k_rescale(const float4* praw, size_t n, const CallScalars* sc, float4* pos)
{
	size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	float4 p = praw[i];
	float scale = sc->scale;
	pos[i] = make_float4((p.x - sc->lo[0]) * scale, (p.y - sc->lo[1]) * scale, (p.z - sc->lo[2]) * scale, 0.f);
}

void launch_rescale(hipStream_t s, Mesh& m)
{
	size_t n = m.vertex_count;
	k_bbox_init<<<1, 1, 0, s>>>(m.sc);
	if (n)
	{
		unsigned int blocks = div_up(n, 256 * 8);
		blocks = blocks > 512 ? 512 : blocks;
		k_bbox<<<blocks, 256, 0, s>>>(m.praw, n, m.sc);
	}
	k_extent<<<1, 1, 0, s>>>(m.praw, m.sc);
	if (n)
		k_rescale<<<grid_for(n), kThreads, 0, s>>>(m.praw, n, m.sc, m.pos);
}

// bounding box and extent only (meshopt_simplifyScale, src/simplifier.cpp:2938-2948)
void launch_extent_only(hipStream_t s, Mesh& m)
{
	size_t n = m.vertex_count;
	k_bbox_init<<<1, 1, 0, s>>>(m.sc);
	if (n)
	{
		unsigned int blocks = div_up(n, 256 * 8);
		blocks = blocks > 512 ? 512 : blocks;
		k_bbox<<<blocks, 256, 0, s>>>(m.praw, n, m.sc);
	}
	k_extent<<<1, 1, 0, s>>>(m.praw, m.sc);
}

// rescaleAttributes, src/simplifier.cpp:605-621 (only attributes with weight > 0)

__global__ void k_rescale_attributes(const float* attrs, size_t stride, const uint32_t* sparse_remap, size_t n, uint32_t count, AttrMap map, float* va)
{
	size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	size_t ri = sparse_remap ? sparse_remap[i] : i;
	const float* a = attrs + ri * stride;
	for (uint32_t k = 0; k < count; ++k)
		va[i * count + k] = a[map.index[k]] * map.weight[k];
}

void launch_rescale_attributes(hipStream_t s, Mesh& m, const float* attributes, size_t stride_floats, const uint32_t* sparse_remap,
                               const uint32_t* attr_index, const float* attr_weight)
{
	if (!m.vertex_count || !m.attr_count)
		return;
	AttrMap map;
	for (uint32_t k = 0; k < m.attr_count; ++k)
	{
		map.index[k] = attr_index[k];
		map.weight[k] = attr_weight[k];
	}
	k_rescale_attributes<<<grid_for(m.vertex_count), kThreads, 0, s>>>(attributes, stride_floats, sparse_remap, m.vertex_count, m.attr_count, map, m.va);
}

// ------------------------------------------------------------------------------------------
// vertex quadrics (src/simplifier.cpp:1057-1157).  One thread per primary r walks r's
// corners in (triangle, corner) order and adds, as the reference's three sequential fills
// do: every face quadric, then the regularising point quadric, then the border/seam edge
// quadrics in (triangle, edge, endpoint) order.  No float atomics; each sum starts at +0.

__device__ __forceinline__ bool edge_loop_qualifies(const uint8_t* kind, const uint32_t* loop, const uint32_t* loopback, uint32_t i0, uint32_t i1, float& weight)
{
	uint8_t k0 = kind[i0], k1 = kind[i1];
	bool b0 = k0 == kBorder || k0 == kSeam;
	bool b1 = k1 == kBorder || k1 == kSeam;
	if (!b0 && !b1)
		return false;
	if (b0 && loop[i0] != i1)
		return false;
	if (b1 && loopback[i1] != i0)
		return false;
	weight = (k0 == kBorder || k1 == kBorder) ? 10.f : 0.5f;
	return true;
}

__global__ void k_vertex_quadrics(const uint32_t* ib, const float4* pos, const uint32_t* remap, const uint8_t* kind, const uint32_t* loop,
                                  const uint32_t* loopback, const uint32_t* woff, const uint32_t* wcorners, const uint8_t* vertex_lock,
                                  const uint32_t* sparse_remap, size_t n, float factor, Quad* vq, Grad* vg)
{
	size_t ri = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (ri >= n)
		return;
	uint32_t r = (uint32_t)ri;
	if (remap[r] != r)
		return;

	Quad Q;
	quad_zero(Q);
	Grad GV = {0.f, 0.f, 0.f, 0.f};
	uint32_t begin = woff[r], end = woff[r + 1];

	for (uint32_t p = begin; p < end; ++p)
	{
		uint32_t t = wcorners[p] / 3;
		V3 p0 = load_pos(pos, ib[t * 3 + 0]), p1 = load_pos(pos, ib[t * 3 + 1]), p2 = load_pos(pos, ib[t * 3 + 2]);
		Quad F;
		quad_from_triangle(F, p0, p1, p2, 1.f);
		quad_add(Q, F);
		if (vg)
		{
			// volume gradients of the position solve, summed in the same corner order
			Grad G;
			grad_volume(G, p0, p1, p2);
			grad_add(GV, G);
		}
	}
	if (vg)
		vg[r] = GV;

	{
		size_t rr = sparse_remap ? sparse_remap[r] : r;
		bool priority = vertex_lock && (vertex_lock[rr] & kVertexPriority) != 0;
		V3 v = load_pos(pos, r);
		float w = Q.w * (priority ? 1.0f : factor);
		Quad P;
		quad_from_point(P, v.x, v.y, v.z, w);
		quad_add(Q, P);
	}

	uint32_t last_t = kNone;
	for (uint32_t p = begin; p < end; ++p)
	{
		uint32_t t = wcorners[p] / 3;
		if (t == last_t)
			continue;
		last_t = t;
		uint32_t tri[3] = {ib[t * 3 + 0], ib[t * 3 + 1], ib[t * 3 + 2]};
		for (int e = 0; e < 3; ++e)
		{
			uint32_t i0 = tri[e];
			uint32_t i1 = tri[e == 2 ? 0 : e + 1];
			uint32_t i2 = tri[e == 0 ? 2 : e - 1];
			bool at0 = remap[i0] == r, at1 = remap[i1] == r;
			if (!at0 && !at1)
				continue;
			float weight;
			if (!edge_loop_qualifies(kind, loop, loopback, i0, i1, weight))
				continue;
			Quad E;
			quad_edge_term(E, load_pos(pos, i0), load_pos(pos, i1), load_pos(pos, i2), weight);
			if (at0)
				quad_add(Q, E);
			if (at1)
				quad_add(Q, E);
		}
	}

	store_quad(vq, r, Q);
}

void launch_vertex_quadrics(hipStream_t s, Mesh& m, const uint32_t* welded_off, const uint32_t* welded_corners, const uint8_t* vertex_lock,
                            const uint32_t* sparse_remap, uint32_t options)
{
	size_t n = m.vertex_count;
	if (!n)
		return;
	float factor = (options & kOptRegularizeLight) ? 1e-2f : ((options & kOptRegularize) ? 1e-1f : 1e-7f);
	k_vertex_quadrics<<<grid_for(n), kThreads, 0, s>>>(m.ib[0], m.pos, m.remap, m.kind, m.loop[0], m.loopback[0], welded_off, welded_corners,
	                                                   vertex_lock, sparse_remap, n, factor, m.vq, m.vg);
}

// attribute quadrics and gradients (src/simplifier.cpp:1159-1179), per wedge (not welded),
// triangles in order.  N is the attribute count when it is small enough for registers.
template <int N>
__global__ void k_attribute_quadrics(const uint32_t* ib, const float4* pos, const float* va, const uint32_t* off, const uint32_t* corners,
                                     size_t n, int count, Quad* aq, Grad* ag)
{
	size_t ii = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (ii >= n)
		return;
	uint32_t i = (uint32_t)ii;
	const int A = N > 0 ? N : count;

	Quad Q;
	quad_zero(Q);
	Grad G[N > 0 ? N : 1];
	if (N > 0)
		for (int k = 0; k < A; ++k)
			G[k] = Grad{0.f, 0.f, 0.f, 0.f};
	else
		for (int k = 0; k < A; ++k)
			ag[(size_t)i * A + k] = Grad{0.f, 0.f, 0.f, 0.f};

	for (uint32_t p = off[i], e = off[i + 1]; p < e; ++p)
	{
		uint32_t t = corners[p] / 3;
		uint32_t i0 = ib[t * 3 + 0], i1 = ib[t * 3 + 1], i2 = ib[t * 3 + 2];
		AttrFrame f = attr_frame(load_pos(pos, i0), load_pos(pos, i1), load_pos(pos, i2));

		Quad QA;
		quad_zero(QA);
		QA.w = f.w;
		const float* a0 = va + (size_t)i0 * A;
		const float* a1 = va + (size_t)i1 * A;
		const float* a2 = va + (size_t)i2 * A;
		for (int k = 0; k < A; ++k)
		{
			Grad g;
			attr_term(QA, g, f, a0[k], a1[k], a2[k]);
			if (N > 0)
				grad_add(G[k], g);
			else
			{
				Grad acc = ag[(size_t)i * A + k];
				grad_add(acc, g);
				ag[(size_t)i * A + k] = acc;
			}
		}
		quad_add(Q, QA);
	}

	store_quad(aq, i, Q);
	if (N > 0)
		for (int k = 0; k < A; ++k)
			ag[(size_t)i * A + k] = G[k];
}

void launch_attribute_quadrics(hipStream_t s, Mesh& m, const uint32_t* raw_off, const uint32_t* raw_corners)
{
	size_t n = m.vertex_count;
	int A = (int)m.attr_count;
	if (!n || !A)
		return;
	unsigned int g = grid_for(n);
	switch (A)
	{
#define CASE(K) \
	case K: \
		k_attribute_quadrics<K><<<g, kThreads, 0, s>>>(m.ib[0], m.pos, m.va, raw_off, raw_corners, n, A, m.aq, m.ag); \
		break;
		CASE(1)
		CASE(2)
		CASE(3)
		CASE(4)
		CASE(5)
		CASE(6)
		CASE(7)
		CASE(8)
#undef CASE
	default:
		k_attribute_quadrics<0><<<g, kThreads, 0, s>>>(m.ib[0], m.pos, m.va, raw_off, raw_corners, n, A, m.aq, m.ag);
	}
}

// boundEdgeCollapses dual count, src/simplifier.cpp:1292-1308
__global__ void __launch_bounds__(256) k_dual_count(const uint8_t* kind, const uint32_t* off, size_t n, uint32_t* out)
{
	__shared__ uint32_t red[4];
	uint32_t v = 0;
	for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
	{
		uint8_t k = kind[i];
		v += (k == kManifold || k == kSeam) ? off[i + 1] - off[i] : 0u;
	}
	for (int o = 32; o > 0; o >>= 1)
		v += __shfl_xor(v, o, 64);
	if (__lane_id() == 0)
		red[threadIdx.x >> 6] = v;
	__syncthreads();
	if (threadIdx.x == 0)
	{
		uint32_t t = red[0] + red[1] + red[2] + red[3];
		if (t)
			atomicAdd(out, t);
	}
}

void launch_dual_count(hipStream_t s, Mesh& m, const uint32_t* raw_off)
{
	HIP_CHECK(hipMemsetAsync(&m.sc->dual_count, 0, sizeof(uint32_t), s));
	if (m.vertex_count)
	{
		unsigned int blocks = div_up(m.vertex_count, 256 * 8);
		blocks = blocks > 512 ? 512 : blocks;
		k_dual_count<<<blocks, 256, 0, s>>>(m.kind, raw_off, m.vertex_count, &m.sc->dual_count);
	}
}

} // namespace mhip
