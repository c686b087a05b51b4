// The Prune option (src/simplifier.cpp:1866-2024): connected components of the welded mesh,
// each component's error (the squared radius of its vertices around their mean), and the removal
// of whole components whose error is at or below a cutoff.
//
// Components come from a lock-free union-find that always hooks the larger root under the
// smaller, so every root is the lowest primary of its component, exactly as the reference's
// sequential union (buildComponents) leaves it; ids are then numbered in root order.  The error
// sums run sequentially per component over its vertices in index order (one thread each), which
// is the reference's accumulation order, so the values match bit for bit.
#include "kernels.h"
#include "prims.h"

namespace mhip
{

static constexpr int kPruneThreads = 256;

static inline unsigned int prune_grid(size_t n)
{
	return div_up(n ? n : 1, kPruneThreads);
}

// root of x with path halving; parent[v] <= v always holds, so a stale read still yields an
// ancestor and every concurrent halving store keeps the trees valid
__device__ __forceinline__ uint32_t uf_root(uint32_t* parent, uint32_t x)
{
	for (;;)
	{
		uint32_t p = __atomic_load_n(&parent[x], __ATOMIC_RELAXED);
		if (p == x)
			return x;
		uint32_t g = __atomic_load_n(&parent[p], __ATOMIC_RELAXED);
		if (g != p)
			__atomic_store_n(&parent[x], g, __ATOMIC_RELAXED);
		x = g;
	}
}

__device__ __forceinline__ void uf_union(uint32_t* parent, uint32_t a, uint32_t b)
{
	for (;;)
	{
		uint32_t ra = uf_root(parent, a), rb = uf_root(parent, b);
		if (ra == rb)
			return;
		uint32_t hi = ra > rb ? ra : rb, lo = ra > rb ? rb : ra;
		if (atomicCAS(&parent[hi], hi, lo) == hi)
			return;
		// hi was hooked by another thread meanwhile: retry from the new roots
	}
}

__global__ void k_uf_init(uint32_t* parent, size_t n)
{
	size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (v < n)
		parent[v] = (uint32_t)v;
}

// buildComponents (src/simplifier.cpp:1866-1908): union over the welded edges of every triangle
__global__ void k_uf_edges(const uint32_t* ib, size_t tris, const uint32_t* remap, uint32_t* parent)
{
	size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= tris)
		return;
	uint32_t a = remap[ib[t * 3 + 0]], b = remap[ib[t * 3 + 1]], c = remap[ib[t * 3 + 2]];
	uf_union(parent, a, b);
	uf_union(parent, b, c);
	uf_union(parent, c, a);
}

__global__ void k_uf_roots(const uint32_t* remap, size_t n, uint32_t* parent, uint32_t* roots, uint32_t* flags)
{
	size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (v >= n)
		return;
	bool primary = remap[v] == (uint32_t)v;
	uint32_t r = primary ? uf_root(parent, (uint32_t)v) : 0;
	roots[v] = r;
	flags[v] = primary && r == (uint32_t)v;
}

__global__ void k_component_ids(const uint32_t* remap, size_t n, const uint32_t* roots, const uint32_t* root_ids, uint32_t* comp)
{
	size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (v >= n)
		return;
	comp[v] = root_ids[roots[remap[v]]];
}

void launch_components(hipStream_t s, Mesh& m, size_t index_count, uint32_t* comp, uint32_t* scratch[5])
{
	const size_t V = m.vertex_count, T = index_count / 3;
	HIP_CHECK(hipMemsetAsync(&m.sc->component_count, 0, sizeof(uint32_t), s));
	if (!V)
		return;
	uint32_t* parent = scratch[0];
	uint32_t* roots = scratch[1];
	uint32_t* ids = scratch[2];
	k_uf_init<<<prune_grid(V), kPruneThreads, 0, s>>>(parent, V);
	if (T)
		k_uf_edges<<<prune_grid(T), kPruneThreads, 0, s>>>(m.ib[0], T, m.remap, parent);
	k_uf_roots<<<prune_grid(V), kPruneThreads, 0, s>>>(m.remap, V, parent, roots, ids);
	scan_exclusive(s, ids, ids, V, nullptr, &m.sc->component_count, m.temp);
	k_component_ids<<<prune_grid(V), kPruneThreads, 0, s>>>(m.remap, V, roots, ids, comp);
}

__global__ void k_copy_u32(const uint32_t* in, uint32_t* out, size_t n)
{
	size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		out[i] = in[i];
}

// start of every component's run in the component-sorted vertex list (every component has its
// root, so no run is empty); off[count] = n
__global__ void k_component_offsets(const uint32_t* sorted_comp, size_t n, uint32_t* off)
{
	size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	uint32_t c = sorted_comp[i];
	if (i == 0 || sorted_comp[i - 1] != c)
		off[c] = (uint32_t)i;
	if (i + 1 == n)
		off[c + 1] = (uint32_t)n;
}

// measureComponents (src/simplifier.cpp:1910-1968): mean, then the largest squared distance
__global__ void k_measure_components(const uint32_t* off, const uint32_t* verts, const float4* pos, const uint32_t* count_dev, float* cerr,
                                     uint32_t* min_bits)
{
	size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (c >= *count_dev)
		return;
	const uint32_t b = off[c], e = off[c + 1];
	float sx = 0.f, sy = 0.f, sz = 0.f, w = 0.f;
	for (uint32_t i = b; i < e; ++i)
	{
		float4 p = pos[verts[i]];
		sx += p.x;
		sy += p.y;
		sz += p.z;
		w += 1.f;
	}
	float iw = w == 0.f ? 0.f : 1.f / w;
	sx *= iw;
	sy *= iw;
	sz *= iw;
	float r2 = 0.f;
	for (uint32_t i = b; i < e; ++i)
	{
		float4 p = pos[verts[i]];
		float dx = p.x - sx, dy = p.y - sy, dz = p.z - sz;
		float r = dx * dx + dy * dy + dz * dz;
		r2 = r2 < r ? r : r2;
	}
	cerr[c] = r2;
	// non-negative floats order like their bit patterns
	atomicMin(min_bits, __float_as_uint(r2));
}

__global__ void k_set_u32(uint32_t* p, uint32_t value)
{
	*p = value;
}

static int bits_for(size_t n)
{
	int b = 1;
	while (b < 32 && (size_t(1) << b) < n)
		++b;
	return b;
}

void launch_measure_components(hipStream_t s, Mesh& m, const uint32_t* comp, float* cerr, uint32_t* scratch[5])
{
	const size_t V = m.vertex_count;
	k_set_u32<<<1, 1, 0, s>>>(&m.sc->prune_min_bits, 0x7f7fffffu); // FLT_MAX
	if (!V)
		return;
	uint32_t* keys = scratch[0];
	uint32_t* vals = scratch[1];
	uint32_t* keys_alt = scratch[2];
	uint32_t* vals_alt = scratch[3];
	uint32_t* off = scratch[4];
	k_copy_u32<<<prune_grid(V), kPruneThreads, 0, s>>>(comp, keys, V);
	iota_u32(s, vals, V);
	// stable: each component's vertices stay in index order (the reference's summation order)
	radix_sort_pairs(s, keys, vals, keys_alt, vals_alt, V, nullptr, bits_for(V), m.temp);
	k_component_offsets<<<prune_grid(V), kPruneThreads, 0, s>>>(keys, V, off);
	k_measure_components<<<prune_grid(V), kPruneThreads, 0, s>>>(off, vals, m.pos, &m.sc->component_count, cerr, &m.sc->prune_min_bits);
}

// pruneComponents (src/simplifier.cpp:1970-2024)
__global__ void k_prune_keep(const uint32_t* ib, size_t tris, const uint32_t* comp, const float* cerr, float cutoff, uint32_t* keep,
                             uint32_t* min_bits)
{
	size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	// the smallest kept error (non-negative floats order like their bits): one atomic per
	// wavefront, so every lane stays to the reduction
	uint32_t bits = 0xffffffffu;
	if (t < tris)
	{
		float e = cerr[comp[ib[t * 3]]];
		bool k = e > cutoff;
		keep[t] = k;
		if (k)
			bits = __float_as_uint(e);
	}
#pragma unroll
	for (int o = 32; o > 0; o >>= 1)
	{
		uint32_t x = __shfl_xor(bits, o, 64);
		bits = x < bits ? x : bits;
	}
	if (__lane_id() == 0 && bits != 0xffffffffu)
		atomicMin(min_bits, bits);
}

__global__ void k_prune_write(const uint32_t* ib, size_t tris, const uint32_t* keep, const uint32_t* off, uint32_t* out)
{
	size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= tris || !keep[t])
		return;
	size_t o = (size_t)off[t] * 3;
	out[o + 0] = ib[t * 3 + 0];
	out[o + 1] = ib[t * 3 + 1];
	out[o + 2] = ib[t * 3 + 2];
}

void launch_prune_triangles(hipStream_t s, Mesh& m, size_t tri_count, const uint32_t* comp, const float* cerr, float cutoff)
{
	HIP_CHECK(hipMemsetAsync(&m.sc->next_triangles, 0, sizeof(uint32_t), s));
	k_set_u32<<<1, 1, 0, s>>>(&m.sc->prune_min_bits, 0x7f7fffffu);
	if (!tri_count)
		return;
	const uint32_t* in = m.ib[m.ib_cur];
	uint32_t* out = m.ib[m.ib_cur ^ 1];
	k_prune_keep<<<prune_grid(tri_count), kPruneThreads, 0, s>>>(in, tri_count, comp, cerr, cutoff, m.tri_cnt, &m.sc->prune_min_bits);
	// keep flags are read by the scatter after the scan overwrites tri_off only
	scan_exclusive(s, m.tri_cnt, m.tri_off, tri_count, nullptr, &m.sc->next_triangles, m.temp);
	k_prune_write<<<prune_grid(tri_count), kPruneThreads, 0, s>>>(in, tri_count, m.tri_cnt, m.tri_off, out);
	m.ib_cur ^= 1;
}

__global__ void k_component_max(const float* cerr, const uint32_t* count_dev, float cutoff, uint32_t* max_bits)
{
	size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (c >= *count_dev)
		return;
	float e = cerr[c];
	if (e > 0.f && e <= cutoff)
		atomicMax(max_bits, __float_as_uint(e));
}

void launch_component_max(hipStream_t s, Mesh& m, const float* cerr, float cutoff)
{
	HIP_CHECK(hipMemsetAsync(&m.sc->prune_max_bits, 0, sizeof(uint32_t), s));
	if (m.vertex_count)
		k_component_max<<<prune_grid(m.vertex_count), kPruneThreads, 0, s>>>(cerr, &m.sc->component_count, cutoff, &m.sc->prune_max_bits);
}

} // namespace mhip
