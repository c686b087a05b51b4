// One clusterlod level on the device: boundary locks, first-use packing of the groups and the
// mapping of the batched results back to shared vertex ids (the simplification itself is the
// batched kernel, batch.hip).  Every per-group kernel runs one workgroup per group; phases that
// read what another thread wrote are separate launches.
#include "clod.h"

namespace mhip
{

static constexpr int kClodThreads = 256;

static inline unsigned int clod_grid(size_t n)
{
	return div_up(n ? n : 1, kClodThreads);
}

// ------------------------------------------------------------------------------------------
// first-use packing: group g hashes its vertices into 2 * (its index count) slots, keeping the
// first position of each; a vertex's packed id is the number of first uses before its own, so the
// packed order is the order Sparse gives (src/simplifier.cpp:151-199).  Each first use also
// records its group in first_group / last_group (lowest and highest group using the remapped
// vertex), which gives lockBoundary's rule (demo/clusterlod.h:372-419) without another pass: a
// remapped vertex is locked when two or more groups use it, i.e. when the two differ.

__device__ __forceinline__ uint32_t clod_slot0(uint32_t v, uint32_t slots)
{
	return (uint32_t)(((uint64_t)(v * 2654435761u) * slots) >> 32);
}

// the slot of vertex v, which the insert pass put into the slice (the probe count is bounded by
// the slice, so no input makes it spin; the host has checked every index against vertex_count)
__device__ __forceinline__ uint32_t clod_find(const uint32_t* key, uint32_t v, uint32_t slots)
{
	uint32_t slot = clod_slot0(v, slots);
	for (uint32_t probe = 0; probe < slots && key[slot] != v + 1; ++probe)
		slot = slot + 1 == slots ? 0 : slot + 1;
	return slot;
}

// insert v at its first position p: at most (index count) distinct keys in twice as many slots,
// so a free or matching slot exists within one sweep of the slice
__device__ __forceinline__ void clod_insert(uint32_t* key, uint32_t* first, uint32_t v, uint32_t p, uint32_t slots)
{
	uint32_t slot = clod_slot0(v, slots);
	for (uint32_t probe = 0; probe < slots; ++probe)
	{
		const uint32_t k = atomicCAS(&key[slot], 0u, v + 1);
		if (k == 0 || k == v + 1)
		{
			atomicMin(&first[slot], p);
			return;
		}
		slot = slot + 1 == slots ? 0 : slot + 1;
	}
}

// exclusive scan of one value per thread over an NT-thread workgroup; lds holds NT / 64 words
template <int NT = kClodThreads>
__device__ __forceinline__ uint32_t clod_block_scan(uint32_t v, uint32_t* lds, uint32_t& total)
{
	uint32_t incl = v;
	for (int o = 1; o < 64; o <<= 1)
	{
		uint32_t t = __shfl_up(incl, o, 64);
		if (__lane_id() >= (uint32_t)o)
			incl += t;
	}
	const uint32_t wave = threadIdx.x / 64;
	if (__lane_id() == 63)
		lds[wave] = incl;
	__syncthreads();
	uint32_t before = 0, all = 0;
	for (uint32_t w = 0; w < NT / 64; ++w)
	{
		before += w < wave ? lds[w] : 0;
		all += lds[w];
	}
	__syncthreads();
	total = all;
	return before + incl - v;
}

__device__ __forceinline__ void clod_first_use(const ClodBuffers& b, uint32_t g, uint32_t at, uint32_t v)
{
	b.staged[at] = v;
	const uint32_t r = b.remap[v];
	atomicMin(&b.first_group[r], g);
	atomicMax(&b.last_group[r], g);
}

static constexpr uint32_t kRanked = 0x80000000u;

// groups of at most lds_slots slots: the whole packing of the group in LDS.  The ordered pass
// replaces a ranked vertex's first position by kRanked | id (positions stay below 2^31), so later
// uses read the id from the same slot and no position compares equal to a ranked entry.  NT
// threads: 256 while several workgroups share a CU's LDS, 1024 when one workgroup takes most of it
// (more wavefronts to hide the LDS atomics and the global loads).
template <int NT>
__global__ void __launch_bounds__(NT) k_clod_group(ClodBuffers b)
{
	extern __shared__ uint32_t lhash[]; // lds_slots keys, then lds_slots first positions
	__shared__ uint32_t lds[NT / 64];
	uint32_t* lkey = lhash;
	uint32_t* lfirst = lhash + b.lds_slots;
	const uint32_t g = blockIdx.x;
	const uint32_t begin = b.index_offsets[g], end = b.index_offsets[g + 1];
	const uint32_t slots = 2 * (end - begin);
	if (slots > b.lds_slots)
		return;
	for (uint32_t q = threadIdx.x; q < slots; q += NT)
	{
		lkey[q] = 0;
		lfirst[q] = kNone;
	}
	__syncthreads();
	for (uint32_t j = begin + threadIdx.x; j < end; j += NT)
		clod_insert(lkey, lfirst, b.indices[j], j - begin, slots);
	__syncthreads();
	uint32_t carry = 0;
	for (uint32_t j0 = begin; j0 < end; j0 += NT)
	{
		const uint32_t j = j0 + threadIdx.x;
		uint32_t v = 0, slot = 0;
		bool is_first = false;
		if (j < end)
		{
			v = b.indices[j];
			slot = clod_find(lkey, v, slots);
			is_first = lfirst[slot] == j - begin;
		}
		uint32_t total;
		const uint32_t id = carry + clod_block_scan<NT>(is_first ? 1u : 0u, lds, total);
		if (is_first)
		{
			lfirst[slot] = kRanked | id;
			clod_first_use(b, g, begin + id, v);
		}
		__syncthreads();
		if (j < end)
			b.packed_indices[j] = lfirst[slot] & ~kRanked;
		carry += total;
	}
	if (threadIdx.x == 0)
		b.unique[g] = carry;
}

// larger groups (beyond the batched kernel's caps; they take the plain path) hash in their global
// slices at 2 * index_offsets[g]: insert, ordered ranking, ids per index as three launches
__global__ void __launch_bounds__(kClodThreads) k_clod_insert(ClodBuffers b)
{
	const uint32_t g = blockIdx.x;
	const uint32_t begin = b.index_offsets[g], end = b.index_offsets[g + 1];
	const uint32_t slots = 2 * (end - begin);
	if (slots <= b.lds_slots)
		return;
	uint32_t* key = b.slot_key + 2 * (size_t)begin;
	uint32_t* first = b.slot_first + 2 * (size_t)begin;
	for (uint32_t j = begin + threadIdx.x; j < end; j += kClodThreads)
		clod_insert(key, first, b.indices[j], j - begin, slots);
}

__global__ void __launch_bounds__(kClodThreads) k_clod_rank(ClodBuffers b)
{
	__shared__ uint32_t lds[kClodThreads / 64];
	const uint32_t g = blockIdx.x;
	const uint32_t begin = b.index_offsets[g], end = b.index_offsets[g + 1];
	const uint32_t slots = 2 * (end - begin);
	if (slots <= b.lds_slots)
		return;
	const uint32_t* key = b.slot_key + 2 * (size_t)begin;
	const uint32_t* first = b.slot_first + 2 * (size_t)begin;
	uint32_t* local = b.slot_local + 2 * (size_t)begin;
	uint32_t carry = 0;
	for (uint32_t j0 = begin; j0 < end; j0 += kClodThreads)
	{
		const uint32_t j = j0 + threadIdx.x;
		uint32_t v = 0, slot = 0;
		bool is_first = false;
		if (j < end)
		{
			v = b.indices[j];
			slot = clod_find(key, v, slots);
			is_first = first[slot] == j - begin;
		}
		uint32_t total;
		const uint32_t id = carry + clod_block_scan(is_first ? 1u : 0u, lds, total);
		if (is_first)
		{
			local[slot] = id;
			clod_first_use(b, g, begin + id, v);
		}
		carry += total;
	}
	if (threadIdx.x == 0)
		b.unique[g] = carry;
}

__global__ void __launch_bounds__(kClodThreads) k_clod_local(ClodBuffers b)
{
	const uint32_t g = blockIdx.x;
	const uint32_t begin = b.index_offsets[g], end = b.index_offsets[g + 1];
	const uint32_t slots = 2 * (end - begin);
	if (slots <= b.lds_slots)
		return;
	const uint32_t* key = b.slot_key + 2 * (size_t)begin;
	const uint32_t* local = b.slot_local + 2 * (size_t)begin;
	for (uint32_t j = begin + threadIdx.x; j < end; j += kClodThreads)
		b.packed_indices[j] = local[clod_find(key, b.indices[j], slots)];
}

uint32_t clod_lds_slots(uint32_t max_group_indices)
{
	// beyond 64 KB of dynamic LDS the kernel needs the runtime's permission
	static const uint32_t cap = [] {
		if (hipFuncSetAttribute(reinterpret_cast<const void*>(&k_clod_group<kClodThreads>), hipFuncAttributeMaxDynamicSharedMemorySize,
		                        (int)(kClodLdsSlotsMax * 8)) == hipSuccess &&
		    hipFuncSetAttribute(reinterpret_cast<const void*>(&k_clod_group<4 * kClodThreads>), hipFuncAttributeMaxDynamicSharedMemorySize,
		                        (int)(kClodLdsSlotsMax * 8)) == hipSuccess)
			return kClodLdsSlotsMax;
		(void)hipGetLastError();
		return 8000u;
	}();
	const uint32_t need = (2 * max_group_indices + 63) & ~63u;
	return need < 64 ? 64 : need < cap ? need : cap;
}

void launch_clod_group(hipStream_t s, const ClodBuffers& b, bool large_groups)
{
	if (!b.group_count)
		return;
	// above 32 KB of hash (4096 slots) fewer than five workgroups share a CU: the wide form
	if (b.lds_slots > 4096)
		k_clod_group<4 * kClodThreads><<<b.group_count, 4 * kClodThreads, 8 * (size_t)b.lds_slots, s>>>(b);
	else
		k_clod_group<kClodThreads><<<b.group_count, kClodThreads, 8 * (size_t)b.lds_slots, s>>>(b);
	if (large_groups)
	{
		k_clod_insert<<<b.group_count, kClodThreads, 0, s>>>(b);
		k_clod_rank<<<b.group_count, kClodThreads, 0, s>>>(b);
		k_clod_local<<<b.group_count, kClodThreads, 0, s>>>(b);
	}
}

// lockBoundary's last loop (demo/clusterlod.h:410-418) runs in index order with remap[i] <= i, so
// when vertex i reads locks[r] & Lock for r < i, vertex r's own caller bits are already in it:
// vertex i gets the boundary bit of its position, the caller's Lock bit of the position's first
// vertex, and its own caller bits
__global__ void k_clod_locks(const uint32_t* first_group, const uint32_t* last_group, const uint32_t* remap, const uint8_t* base, size_t n,
                             uint8_t* locks)
{
	size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
	{
		const uint32_t r = remap[i], f = first_group[r];
		const uint8_t boundary = f != kNone && f != last_group[r] ? kVertexLock : 0;
		locks[i] = boundary | (base ? (uint8_t)((base[r] & kVertexLock) | base[i]) : 0);
	}
}

void launch_clod_locks(hipStream_t s, const ClodBuffers& b, const uint8_t* base_locks, uint8_t* locks)
{
	if (b.vertex_count)
		k_clod_locks<<<clod_grid(b.vertex_count), kClodThreads, 0, s>>>(b.first_group, b.last_group, b.remap, base_locks, b.vertex_count, locks);
}

// group g's packed vertices at unique[g] (the scanned counts): shared id, position, lock and
// attributes
__global__ void __launch_bounds__(kClodThreads) k_clod_place(ClodBuffers b, const float* positions, size_t stride, const uint8_t* locks,
                                                             const float* attributes, size_t attribute_stride)
{
	const uint32_t g = blockIdx.x;
	const uint32_t begin = b.index_offsets[g], vbase = b.unique[g], n = b.unique[g + 1] - vbase;
	const uint32_t A = b.attribute_count;
	for (uint32_t k = threadIdx.x; k < n; k += kClodThreads)
	{
		const uint32_t v = b.staged[begin + k], at = vbase + k;
		b.order[at] = v;
		const float* p = positions + (size_t)v * stride;
		b.packed_positions[3 * (size_t)at + 0] = p[0];
		b.packed_positions[3 * (size_t)at + 1] = p[1];
		b.packed_positions[3 * (size_t)at + 2] = p[2];
		b.packed_locks[at] = locks[v];
		for (uint32_t q = 0; q < A; ++q)
			b.packed_attributes[(size_t)at * A + q] = attributes[(size_t)v * attribute_stride + q];
	}
}

void launch_clod_place(hipStream_t s, const ClodBuffers& b, const float* positions, size_t stride_floats, const uint8_t* locks,
                       const float* attributes, size_t attribute_stride_floats)
{
	if (b.group_count)
		k_clod_place<<<b.group_count, kClodThreads, 0, s>>>(b, positions, stride_floats, locks, attributes, attribute_stride_floats);
}

__global__ void __launch_bounds__(kClodThreads) k_clod_unpack(ClodBuffers b, const uint32_t* counts, uint32_t* destination)
{
	const uint32_t g = blockIdx.x;
	const uint32_t begin = b.index_offsets[g], n = counts[g];
	const uint32_t* order = b.order + b.unique[g];
	for (uint32_t k = threadIdx.x; k < n; k += kClodThreads)
		destination[begin + k] = order[b.packed_result[begin + k]];
}

void launch_clod_unpack(hipStream_t s, const ClodBuffers& b, const uint32_t* counts, uint32_t* destination)
{
	if (b.group_count)
		k_clod_unpack<<<b.group_count, kClodThreads, 0, s>>>(b, counts, destination);
}

// ------------------------------------------------------------------------------------------
// the per-group fallbacks of clodBuild's simplify (demo/clusterlod.h:461-518)

// groups list[f] of a sub-batch: their packed indices, positions and locks gathered at the
// sub-batch's offsets (sub_io, sub_vo, host-computed scans of the groups' sizes)
__global__ void __launch_bounds__(kClodThreads) k_clod_gather(ClodBuffers b, const uint32_t* list, const uint32_t* sub_io, const uint32_t* sub_vo,
                                                              uint32_t* sub_idx, float* sub_pos, uint8_t* sub_locks, float* sub_attr)
{
	const uint32_t A = b.attribute_count;
	const uint32_t f = blockIdx.x, g = list[f];
	const uint32_t ib = b.index_offsets[g], ni = b.index_offsets[g + 1] - ib, vb = b.unique[g], nv = b.unique[g + 1] - vb;
	for (uint32_t k = threadIdx.x; k < ni; k += kClodThreads)
		sub_idx[sub_io[f] + k] = b.packed_indices[ib + k];
	for (uint32_t k = threadIdx.x; k < nv; k += kClodThreads)
	{
		const size_t at = sub_vo[f] + k;
		sub_pos[3 * at + 0] = b.packed_positions[3 * ((size_t)vb + k) + 0];
		sub_pos[3 * at + 1] = b.packed_positions[3 * ((size_t)vb + k) + 1];
		sub_pos[3 * at + 2] = b.packed_positions[3 * ((size_t)vb + k) + 2];
		sub_locks[at] = b.packed_locks[vb + k];
		for (uint32_t q = 0; q < A; ++q)
			sub_attr[at * A + q] = b.packed_attributes[((size_t)vb + k) * A + q];
	}
}

// the sub-batch's results replace those of its groups (a group's packed ids are the same in both)
__global__ void __launch_bounds__(kClodThreads) k_clod_scatter(ClodBuffers b, const uint32_t* list, const uint32_t* sub_io, const uint32_t* sub_result,
                                                               const uint32_t* sub_counts, const float* sub_errors, uint32_t* counts, float* errors)
{
	const uint32_t f = blockIdx.x, g = list[f];
	const uint32_t n = sub_counts[f];
	for (uint32_t k = threadIdx.x; k < n; k += kClodThreads)
		b.packed_result[b.index_offsets[g] + k] = sub_result[sub_io[f] + k];
	if (threadIdx.x == 0)
	{
		counts[g] = n;
		if (errors)
			errors[g] = sub_errors[f];
	}
}

void launch_clod_gather(hipStream_t s, const ClodBuffers& b, const uint32_t* list, uint32_t n, const uint32_t* sub_io, const uint32_t* sub_vo,
                        uint32_t* sub_idx, float* sub_pos, uint8_t* sub_locks, float* sub_attr)
{
	if (n)
		k_clod_gather<<<n, kClodThreads, 0, s>>>(b, list, sub_io, sub_vo, sub_idx, sub_pos, sub_locks, sub_attr);
}

void launch_clod_scatter(hipStream_t s, const ClodBuffers& b, const uint32_t* list, uint32_t n, const uint32_t* sub_io, const uint32_t* sub_result,
                         const uint32_t* sub_counts, const float* sub_errors, uint32_t* counts, float* errors)
{
	if (n)
		k_clod_scatter<<<n, kClodThreads, 0, s>>>(b, list, sub_io, sub_result, sub_counts, sub_errors, counts, errors);
}

// simplifyFallback's subset (demo/clusterlod.h:426-458): one vertex per index of group g (its
// position and lock), and the identity index buffer over them
__global__ void k_clod_deindex(ClodBuffers b, uint32_t g, float* subset_pos, uint8_t* subset_locks, uint32_t* subset_idx)
{
	const uint32_t ib = b.index_offsets[g], n = b.index_offsets[g + 1] - ib, vb = b.unique[g];
	for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
	{
		const size_t v = (size_t)vb + b.packed_indices[ib + i];
		subset_pos[3 * (size_t)i + 0] = b.packed_positions[3 * v + 0];
		subset_pos[3 * (size_t)i + 1] = b.packed_positions[3 * v + 1];
		subset_pos[3 * (size_t)i + 2] = b.packed_positions[3 * v + 2];
		subset_locks[i] = b.packed_locks[v];
		subset_idx[i] = i;
	}
}

// the sloppy result over the subset becomes group g's result: subset vertex i is the group's index i
__global__ void k_clod_sloppy_result(ClodBuffers b, uint32_t g, const uint32_t* lod, uint32_t count, uint32_t* counts, float* errors, float error)
{
	const uint32_t ib = b.index_offsets[g];
	for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < count; k += gridDim.x * blockDim.x)
		b.packed_result[ib + k] = b.packed_indices[ib + lod[k]];
	if (blockIdx.x == 0 && threadIdx.x == 0)
	{
		counts[g] = count;
		if (errors)
			errors[g] = error;
	}
}

void launch_clod_deindex(hipStream_t s, const ClodBuffers& b, uint32_t g, uint32_t n, float* subset_pos, uint8_t* subset_locks, uint32_t* subset_idx)
{
	k_clod_deindex<<<clod_grid(n ? n : 1), kClodThreads, 0, s>>>(b, g, subset_pos, subset_locks, subset_idx);
}

void launch_clod_sloppy_result(hipStream_t s, const ClodBuffers& b, uint32_t g, const uint32_t* lod, uint32_t count, uint32_t* counts, float* errors,
                               float error)
{
	k_clod_sloppy_result<<<clod_grid(count ? count : 1), kClodThreads, 0, s>>>(b, g, lod, count, counts, errors, error);
}

// the edge-length limit (demo/clusterlod.h:489-516): over the group's input triangles, the largest
// max(shortest squared edge, longest / 4) with the std::max / std::min of the reference (a NaN
// never replaces the running maximum, which starts at 0: the result is order-independent), then
// error = min(error, sqrt(that) * limit)
__device__ __forceinline__ float std_max(float a, float b)
{
	return a < b ? b : a;
}

__device__ __forceinline__ float std_min(float a, float b)
{
	return b < a ? b : a;
}

__global__ void __launch_bounds__(kClodThreads) k_clod_edge_limit(ClodBuffers b, const uint32_t* targets, float limit, float* errors)
{
	__shared__ float part[kClodThreads / 64];
	const uint32_t g = blockIdx.x;
	const uint32_t ib = b.index_offsets[g], n = b.index_offsets[g + 1] - ib, vb = b.unique[g];
	// a group whose target exceeds its size keeps its indices and error unchanged (the early return
	// of demo/clusterlod.h:463-464 comes before the limit)
	if (targets[g] > n)
		return;
	float m = 0.f;
	for (uint32_t t = threadIdx.x; t < n / 3; t += kClodThreads)
	{
		const float* va = b.packed_positions + 3 * ((size_t)vb + b.packed_indices[ib + 3 * t + 0]);
		const float* vv = b.packed_positions + 3 * ((size_t)vb + b.packed_indices[ib + 3 * t + 1]);
		const float* vc = b.packed_positions + 3 * ((size_t)vb + b.packed_indices[ib + 3 * t + 2]);
		const float eab = (va[0] - vv[0]) * (va[0] - vv[0]) + (va[1] - vv[1]) * (va[1] - vv[1]) + (va[2] - vv[2]) * (va[2] - vv[2]);
		const float eac = (va[0] - vc[0]) * (va[0] - vc[0]) + (va[1] - vc[1]) * (va[1] - vc[1]) + (va[2] - vc[2]) * (va[2] - vc[2]);
		const float ebc = (vv[0] - vc[0]) * (vv[0] - vc[0]) + (vv[1] - vc[1]) * (vv[1] - vc[1]) + (vv[2] - vc[2]) * (vv[2] - vc[2]);
		const float emax = std_max(std_max(eab, eac), ebc);
		const float emin = std_min(std_min(eab, eac), ebc);
		m = std_max(m, std_max(emin, emax / 4));
	}
	for (int o = 32; o > 0; o >>= 1)
		m = std_max(m, __shfl_xor(m, o, 64));
	if (__lane_id() == 0)
		part[threadIdx.x >> 6] = m;
	__syncthreads();
	if (threadIdx.x == 0)
	{
		float all = 0.f;
		for (int w = 0; w < kClodThreads / 64; ++w)
			all = std_max(all, part[w]);
		errors[g] = std_min(errors[g], sqrtf(all) * limit);
	}
}

void launch_clod_edge_limit(hipStream_t s, const ClodBuffers& b, const uint32_t* targets, float limit, float* errors)
{
	if (b.group_count)
		k_clod_edge_limit<<<b.group_count, kClodThreads, 0, s>>>(b, targets, limit, errors);
}

// groups whose target exceeds their size: the result is the group's own index buffer, error 0
// (demo/clusterlod.h:463-464 returns the indices and leaves the caller's error, 0 in clodBuild)
__global__ void __launch_bounds__(kClodThreads) k_clod_keep(ClodBuffers b, const uint32_t* list, uint32_t* counts, float* errors)
{
	const uint32_t g = list[blockIdx.x];
	const uint32_t ib = b.index_offsets[g], n = b.index_offsets[g + 1] - ib;
	for (uint32_t k = threadIdx.x; k < n; k += kClodThreads)
		b.packed_result[ib + k] = b.packed_indices[ib + k];
	if (threadIdx.x == 0)
	{
		counts[g] = n;
		if (errors)
			errors[g] = 0.f;
	}
}

void launch_clod_keep(hipStream_t s, const ClodBuffers& b, const uint32_t* list, uint32_t n, uint32_t* counts, float* errors)
{
	if (n)
		k_clod_keep<<<n, kClodThreads, 0, s>>>(b, list, counts, errors);
}

} // namespace mhip
