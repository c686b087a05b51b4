// meshopt_simplifySloppy kernels (src/simplifier.cpp:2087-2334, 2647-2778).  Every step that the
// reference runs in input order is made order-independent: cells are numbered by the first
// vertex of each id (a minimum per id, then a scan), cell quadrics are summed per cell in the
// reference's (triangle, corner) order after a stable sort, the chosen vertex is a 64-bit
// minimum of (error bits, index), and a duplicate triangle keeps its first occurrence (a minimum
// per canonical triple, then an ordered compaction).  No float atomics.
#include "prims.h"
#include "quadric.h"
#include "sloppy.h"

namespace mhip
{

static constexpr int kThreads = 256;

static inline unsigned int grid_for(size_t n, unsigned int cap = 65535)
{
	unsigned int g = div_up(n ? n : 1, kThreads);
	return g > cap ? cap : g;
}

// float to int as the reference's x86 build converts it (cvttss2si): truncation toward zero, and
// the "integer indefinite" 0x80000000 for NaN and for values outside the int range.  The GPU's
// own conversion gives 0 for NaN and saturates, so a NaN position (the rescale leaves NaN
// coordinates NaN) would fall into a different cell than on the CPU.
__device__ __forceinline__ int cvt_trunc_x86(float f)
{
	return (f >= -2147483648.f && f < 2147483648.f) ? (int)f : (int)0x80000000u;
}

// computeVertexIds (src/simplifier.cpp:2087-2105): round-half-up cell coordinates of the
// rescaled position, 10 bits per axis; a vertex locked by vertex_lock gets an id of its own
__device__ __forceinline__ uint32_t sloppy_id(const float4* pos, const uint8_t* lock, uint32_t v, float cell_scale)
{
	if (lock && (lock[v] & kVertexLock))
		return (1u << 30) | v;
	const float4 p = pos[v];
	const int xi = cvt_trunc_x86(p.x * cell_scale + 0.5f);
	const int yi = cvt_trunc_x86(p.y * cell_scale + 0.5f);
	const int zi = cvt_trunc_x86(p.z * cell_scale + 0.5f);
	return ((uint32_t)xi << 20) | ((uint32_t)yi << 10) | (uint32_t)zi;
}

__device__ __forceinline__ uint32_t mix32(uint32_t h)
{
	h ^= h >> 16;
	h *= 0x7feb352du;
	h ^= h >> 15;
	h *= 0x846ca68bu;
	h ^= h >> 16;
	return h;
}

// countTriangles (src/simplifier.cpp:2107-2121) at one grid size, ids computed on the fly; one
// atomic per workgroup (same-address atomics serialise in L2, so the grid is kept small)
__global__ void __launch_bounds__(kThreads) k_sloppy_count(const uint32_t* idx, uint32_t tris, const float4* pos, const uint8_t* lock, float cell_scale,
                                                           uint32_t* count)
{
	__shared__ uint32_t part[kThreads / 64];
	uint32_t local = 0;
	for (uint32_t t = blockIdx.x * kThreads + threadIdx.x; t < tris; t += gridDim.x * kThreads)
	{
		const uint32_t a = sloppy_id(pos, lock, idx[3 * t + 0], cell_scale);
		const uint32_t b = sloppy_id(pos, lock, idx[3 * t + 1], cell_scale);
		const uint32_t c = sloppy_id(pos, lock, idx[3 * t + 2], cell_scale);
		local += (a != b) & (a != c) & (b != c);
	}
	for (int o = 32; o > 0; o >>= 1)
		local += __shfl_xor(local, o, 64);
	if (__lane_id() == 0)
		part[threadIdx.x >> 6] = local;
	__syncthreads();
	if (threadIdx.x == 0)
	{
		uint32_t sum = 0;
		for (int w = 0; w < kThreads / 64; ++w)
			sum += part[w];
		if (sum)
			atomicAdd(count, sum);
	}
}

__global__ void k_sloppy_ids(const float4* pos, const uint8_t* lock, uint32_t n, float cell_scale, uint32_t* ids)
{
	for (uint32_t v = blockIdx.x * kThreads + threadIdx.x; v < n; v += gridDim.x * kThreads)
		ids[v] = sloppy_id(pos, lock, v, cell_scale);
}

// fillVertexCells (src/simplifier.cpp:2123-2146), step 1: the table slot of an id ends holding
// the lowest vertex with that id (a slot only ever holds vertices of one id)
__global__ void k_vtable_insert(const uint32_t* ids, uint32_t n, uint32_t* table, uint32_t mask)
{
	for (uint32_t v = blockIdx.x * kThreads + threadIdx.x; v < n; v += gridDim.x * kThreads)
	{
		const uint32_t id = ids[v];
		uint32_t slot = mix32(id) & mask;
		for (uint32_t probe = 0; probe <= mask; ++probe)
		{
			const uint32_t old = atomicCAS(&table[slot], kNone, v);
			if (old == kNone)
				break;
			if (ids[old] == id)
			{
				atomicMin(&table[slot], v);
				break;
			}
			slot = (slot + 1) & mask;
		}
	}
}

// step 2: each vertex's representative (the first vertex of its id) and the representative flags
__global__ void k_vtable_rep(const uint32_t* ids, uint32_t n, const uint32_t* table, uint32_t mask, uint32_t* rep, uint32_t* flag)
{
	for (uint32_t v = blockIdx.x * kThreads + threadIdx.x; v < n; v += gridDim.x * kThreads)
	{
		const uint32_t id = ids[v];
		uint32_t slot = mix32(id) & mask, r = v;
		for (uint32_t probe = 0; probe <= mask; ++probe)
		{
			const uint32_t e = table[slot];
			if (e == kNone)
				break;
			if (ids[e] == id)
			{
				r = e;
				break;
			}
			slot = (slot + 1) & mask;
		}
		rep[v] = r;
		flag[v] = r == v ? 1u : 0u;
	}
}

// step 3: cell numbers are the ranks of the representatives in vertex order
__global__ void k_cells(const uint32_t* rep, const uint32_t* scan, uint32_t n, uint32_t* cells, uint32_t* cell_first)
{
	for (uint32_t v = blockIdx.x * kThreads + threadIdx.x; v < n; v += gridDim.x * kThreads)
	{
		const uint32_t c = scan[rep[v]];
		cells[v] = c;
		if (rep[v] == v)
			cell_first[c] = v;
	}
}

// fillCellQuadrics (src/simplifier.cpp:2169-2199), step 1: one entry 3t + k per corner that adds
// triangle t's quadric to a cell (only corner 0 when the triangle lies in one cell); other
// entries get the key cell_count and sort to the end
__global__ void k_cell_entries(const uint32_t* idx, uint32_t tris, const uint32_t* cells, uint32_t cell_count, uint32_t* keys, uint32_t* vals)
{
	for (uint32_t e = blockIdx.x * kThreads + threadIdx.x; e < 3 * tris; e += gridDim.x * kThreads)
	{
		const uint32_t t = e / 3, k = e - 3 * t;
		const uint32_t c0 = cells[idx[3 * t + 0]], c1 = cells[idx[3 * t + 1]], c2 = cells[idx[3 * t + 2]];
		const bool single = c0 == c1 && c0 == c2;
		const uint32_t ck = k == 0 ? c0 : (k == 1 ? c1 : c2);
		const uint32_t key = (single && k) ? cell_count : ck;
		keys[e] = key;
		vals[e] = e;
	}
}

// step 2: the start of each cell's run in the sorted keys (lower bound), off[cell_count] = the
// number of entries that add to a cell
__global__ void k_cell_runs(const uint32_t* keys, uint32_t n, uint32_t cell_count, uint32_t* off)
{
	for (uint32_t c = blockIdx.x * kThreads + threadIdx.x; c <= cell_count; c += gridDim.x * kThreads)
	{
		uint32_t lo = 0, hi = n;
		while (lo < hi)
		{
			const uint32_t mid = lo + (hi - lo) / 2;
			if (keys[mid] < c)
				lo = mid + 1;
			else
				hi = mid;
		}
		off[c] = lo;
	}
}

// step 3a: the face quadric of every triangle, once (a triangle with corners in two or three
// cells adds the same weight-1 quadric to each; one inside a single cell adds weight 3)
__global__ void k_face_quadrics(const uint32_t* idx, uint32_t tris, const uint32_t* cells, const float4* pos, Quad* face)
{
	const uint32_t t = blockIdx.x * kThreads + threadIdx.x;
	if (t >= tris)
		return;
	const uint32_t i0 = idx[3 * t + 0], i1 = idx[3 * t + 1], i2 = idx[3 * t + 2];
	const uint32_t c0 = cells[i0], c1 = cells[i1], c2 = cells[i2];
	const bool single = c0 == c1 && c0 == c2;
	Quad F;
	quad_from_triangle(F, load_pos(pos, i0), load_pos(pos, i1), load_pos(pos, i2), single ? 3.f : 1.f);
	store_quad(face, t, F);
}

// step 3b: each cell sums its entries' face quadrics from zero in entry order (the stable sort
// kept the reference's (triangle, corner) order).  Four entries' quadrics are loaded before they
// are added one after the other, so the loads overlap and the additions keep the sequential order.
__global__ void k_cell_quadrics(const uint32_t* vals, const uint32_t* off, uint32_t cell_count, const Quad* face, Quad* cq)
{
	for (uint32_t c = blockIdx.x * kThreads + threadIdx.x; c < cell_count; c += gridDim.x * kThreads)
	{
		Quad Q;
		quad_zero(Q);
		uint32_t p = off[c];
		const uint32_t e = off[c + 1];
		for (; p + 4 <= e; p += 4)
		{
			Quad F[4];
#pragma unroll
			for (int u = 0; u < 4; ++u)
				F[u] = load_quad(face, vals[p + u] / 3);
#pragma unroll
			for (int u = 0; u < 4; ++u)
				quad_add(Q, F[u]);
		}
		for (; p < e; ++p)
			quad_add(Q, load_quad(face, vals[p] / 3));
		store_quad(cq, c, Q);
	}
}

// fillCellRemap (src/simplifier.cpp:2237-2254), step 1: the smallest (error, vertex) per cell
// over vertices with a non-NaN error (errors are non-negative, so their bits order like them)
__global__ void k_cell_best(const uint32_t* cells, const Quad* cq, const float4* pos, uint32_t n, unsigned long long* best)
{
	for (uint32_t v = blockIdx.x * kThreads + threadIdx.x; v < n; v += gridDim.x * kThreads)
	{
		const uint32_t c = cells[v];
		const float e = quad_error(load_quad(cq, c), load_pos(pos, v));
		if (e == e)
			atomicMin(&best[c], ((unsigned long long)__float_as_uint(e) << 32) | v);
	}
}

// step 2: the sequential scan keeps the cell's first vertex when its error is NaN (no comparison
// with NaN replaces it), else the first vertex of smallest error; the result error is the
// largest non-NaN cell error
__global__ void k_cell_final(const unsigned long long* best, const uint32_t* cell_first, const Quad* cq, const float4* pos, uint32_t cell_count,
                             uint32_t* cell_remap, uint32_t* error_bits)
{
	for (uint32_t c = blockIdx.x * kThreads + threadIdx.x; c < cell_count; c += gridDim.x * kThreads)
	{
		const uint32_t first = cell_first[c];
		const float e0 = quad_error(load_quad(cq, c), load_pos(pos, first));
		if (e0 != e0)
		{
			cell_remap[c] = first;
			continue;
		}
		const unsigned long long b = best[c];
		cell_remap[c] = (uint32_t)b;
		atomicMax(error_bits, (uint32_t)(b >> 32));
	}
}

// filterTriangles (src/simplifier.cpp:2282-2325), step 1: the canonical triple of each triangle
// over three distinct cells (cell vertices, rotated so the smallest leads)
__global__ void k_tri_canon(const uint32_t* idx, uint32_t tris, const uint32_t* cells, const uint32_t* cell_remap, uint4* tri)
{
	for (uint32_t t = blockIdx.x * kThreads + threadIdx.x; t < tris; t += gridDim.x * kThreads)
	{
		const uint32_t c0 = cells[idx[3 * t + 0]], c1 = cells[idx[3 * t + 1]], c2 = cells[idx[3 * t + 2]];
		uint4 r = make_uint4(0, 0, 0, 0);
		if (c0 != c1 && c0 != c2 && c1 != c2)
		{
			uint32_t a = cell_remap[c0], b = cell_remap[c1], c = cell_remap[c2];
			if (b < a && b < c)
			{
				const uint32_t x = a;
				a = b, b = c, c = x;
			}
			else if (c < a && c < b)
			{
				const uint32_t x = c;
				c = b, b = a, a = x;
			}
			r = make_uint4(a, b, c, 1u);
		}
		tri[t] = r;
	}
}

__device__ __forceinline__ uint32_t tri_hash(uint4 r)
{
	return mix32(r.x * 0x9e3779b1u ^ mix32(r.y) ^ mix32(r.z * 0x85ebca6bu));
}

__device__ __forceinline__ bool same_tri(uint4 a, uint4 b)
{
	return a.x == b.x && a.y == b.y && a.z == b.z;
}

// step 2: the table slot of a triple ends holding its first triangle
__global__ void k_tri_insert(const uint4* tri, uint32_t tris, uint32_t* table, uint32_t mask)
{
	for (uint32_t t = blockIdx.x * kThreads + threadIdx.x; t < tris; t += gridDim.x * kThreads)
	{
		const uint4 r = tri[t];
		if (!r.w)
			continue;
		uint32_t slot = tri_hash(r) & mask;
		for (uint32_t probe = 0; probe <= mask; ++probe)
		{
			const uint32_t old = atomicCAS(&table[slot], kNone, t);
			if (old == kNone)
				break;
			if (same_tri(tri[old], r))
			{
				atomicMin(&table[slot], t);
				break;
			}
			slot = (slot + 1) & mask;
		}
	}
}

// step 3: a triangle is kept when it is the first of its triple
__global__ void k_tri_keep(const uint4* tri, uint32_t tris, const uint32_t* table, uint32_t mask, uint32_t* keep)
{
	for (uint32_t t = blockIdx.x * kThreads + threadIdx.x; t < tris; t += gridDim.x * kThreads)
	{
		const uint4 r = tri[t];
		uint32_t k = 0;
		if (r.w)
		{
			uint32_t slot = tri_hash(r) & mask;
			for (uint32_t probe = 0; probe <= mask; ++probe)
			{
				const uint32_t e = table[slot];
				if (e == kNone)
					break;
				if (same_tri(tri[e], r))
				{
					k = e == t ? 1u : 0u;
					break;
				}
				slot = (slot + 1) & mask;
			}
		}
		keep[t] = k;
	}
}

// step 4: kept triangles in input order
__global__ void k_tri_write(const uint4* tri, const uint32_t* keep, const uint32_t* pos_out, uint32_t tris, uint32_t* destination)
{
	for (uint32_t t = blockIdx.x * kThreads + threadIdx.x; t < tris; t += gridDim.x * kThreads)
	{
		if (!keep[t])
			continue;
		const uint4 r = tri[t];
		const uint32_t o = pos_out[t] * 3;
		destination[o + 0] = r.x;
		destination[o + 1] = r.y;
		destination[o + 2] = r.z;
	}
}

static float cell_scale_of(int grid_size)
{
	return (float)(grid_size - 1);
}

void launch_sloppy_count(hipStream_t s, const SloppyBuffers& b, int grid_size)
{
	const uint32_t tris = b.index_count / 3;
	HIP_CHECK(hipMemsetAsync(b.count, 0, sizeof(uint32_t), s));
	if (tris)
		k_sloppy_count<<<grid_for(tris, 1024), kThreads, 0, s>>>(b.indices, tris, b.pos, b.lock, cell_scale_of(grid_size), b.count);
}

void launch_sloppy_cells(hipStream_t s, const SloppyBuffers& b, int grid_size, uint32_t* cell_count_dev)
{
	const uint32_t n = b.vertex_count;
	k_sloppy_ids<<<grid_for(n), kThreads, 0, s>>>(b.pos, b.lock, n, cell_scale_of(grid_size), b.ids);
	fill_u32(s, b.vtable, kNone, b.vtable_size);
	k_vtable_insert<<<grid_for(n), kThreads, 0, s>>>(b.ids, n, b.vtable, b.vtable_size - 1);
	k_vtable_rep<<<grid_for(n), kThreads, 0, s>>>(b.ids, n, b.vtable, b.vtable_size - 1, b.rep, b.cell_scan);
	scan_exclusive(s, b.cell_scan, b.cell_scan, n, nullptr, cell_count_dev, b.scan_temp);
	k_cells<<<grid_for(n), kThreads, 0, s>>>(b.rep, b.cell_scan, n, b.cells, b.cell_first);
}

static int bits_for(uint32_t max_value)
{
	int bits = 1;
	while (bits < 32 && (max_value >> bits) != 0)
		++bits;
	return bits;
}

void launch_sloppy_quadrics(hipStream_t s, const SloppyBuffers& b, uint32_t cell_count)
{
	const uint32_t tris = b.index_count / 3, n = b.vertex_count;
	if (tris)
	{
		k_cell_entries<<<grid_for(3 * (size_t)tris), kThreads, 0, s>>>(b.indices, tris, b.cells, cell_count, b.entry_keys, b.entry_vals);
		radix_sort_pairs(s, b.entry_keys, b.entry_vals, b.entry_keys_alt, b.entry_vals_alt, 3 * (size_t)tris, nullptr, bits_for(cell_count),
		                 b.sort_temp);
	}
	k_cell_runs<<<grid_for((size_t)cell_count + 1), kThreads, 0, s>>>(b.entry_keys, 3 * tris, cell_count, b.cell_off);
	k_face_quadrics<<<grid_for(tris), kThreads, 0, s>>>(b.indices, tris, b.cells, b.pos, b.face);
	k_cell_quadrics<<<grid_for(cell_count), kThreads, 0, s>>>(b.entry_vals, b.cell_off, cell_count, b.face, b.cq);
	fill_u32(s, reinterpret_cast<uint32_t*>(b.best), kNone, 2 * (size_t)cell_count);
	HIP_CHECK(hipMemsetAsync(b.error_bits, 0, sizeof(uint32_t), s));
	k_cell_best<<<grid_for(n), kThreads, 0, s>>>(b.cells, b.cq, b.pos, n, reinterpret_cast<unsigned long long*>(b.best));
	k_cell_final<<<grid_for(cell_count), kThreads, 0, s>>>(reinterpret_cast<const unsigned long long*>(b.best), b.cell_first, b.cq, b.pos, cell_count,
	                                                       b.cell_remap, b.error_bits);
}

void launch_sloppy_filter(hipStream_t s, const SloppyBuffers& b, uint32_t cell_count, uint32_t* destination, uint32_t* kept_dev)
{
	(void)cell_count;
	const uint32_t tris = b.index_count / 3;
	if (!tris)
	{
		HIP_CHECK(hipMemsetAsync(kept_dev, 0, sizeof(uint32_t), s));
		return;
	}
	k_tri_canon<<<grid_for(tris), kThreads, 0, s>>>(b.indices, tris, b.cells, b.cell_remap, b.tri);
	fill_u32(s, b.ttable, kNone, b.ttable_size);
	k_tri_insert<<<grid_for(tris), kThreads, 0, s>>>(b.tri, tris, b.ttable, b.ttable_size - 1);
	k_tri_keep<<<grid_for(tris), kThreads, 0, s>>>(b.tri, tris, b.ttable, b.ttable_size - 1, b.keep);
	// the entry arrays are free again: the exclusive scan of the keep flags goes to entry_keys
	scan_exclusive(s, b.keep, b.entry_keys, tris, nullptr, kept_dev, b.scan_temp);
	k_tri_write<<<grid_for(tris), kThreads, 0, s>>>(b.tri, b.keep, b.entry_keys, tris, destination);
}

} // namespace mhip
