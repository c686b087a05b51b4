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
// lockBoundary (demo/clusterlod.h:372-419): a remapped vertex is locked when a group other than
// the lowest one that uses it uses it too, i.e. when two or more groups use it

__global__ void __launch_bounds__(kClodThreads) k_clod_first(ClodBuffers b)
{
	const uint32_t g = blockIdx.x;
	for (uint32_t j = b.index_offsets[g] + threadIdx.x, e = b.index_offsets[g + 1]; j < e; j += kClodThreads)
		atomicMin(&b.first_group[b.remap[b.indices[j]]], g);
}

__global__ void __launch_bounds__(kClodThreads) k_clod_boundary(ClodBuffers b)
{
	const uint32_t g = blockIdx.x;
	for (uint32_t j = b.index_offsets[g] + threadIdx.x, e = b.index_offsets[g + 1]; j < e; j += kClodThreads)
	{
		const uint32_t r = b.remap[b.indices[j]];
		if (b.first_group[r] != g)
			b.boundary[r] = 1;
	}
}

__global__ void k_clod_locks(const uint8_t* boundary, const uint32_t* remap, const uint8_t* base, size_t n, uint8_t* locks)
{
	size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		locks[i] = (boundary[remap[i]] ? kVertexLock : 0) | (base ? base[i] : 0);
}

void launch_clod_locks(hipStream_t s, const ClodBuffers& b, const uint8_t* base_locks, uint8_t* locks)
{
	const size_t n = b.vertex_count;
	if (!n)
		return;
	HIP_CHECK(hipMemsetAsync(b.first_group, 0xff, n * sizeof(uint32_t), s));
	HIP_CHECK(hipMemsetAsync(b.boundary, 0, n, s));
	if (b.group_count)
	{
		k_clod_first<<<b.group_count, kClodThreads, 0, s>>>(b);
		k_clod_boundary<<<b.group_count, kClodThreads, 0, s>>>(b);
	}
	k_clod_locks<<<clod_grid(n), kClodThreads, 0, s>>>(b.boundary, b.remap, base_locks, n, locks);
}

// ------------------------------------------------------------------------------------------
// first-use packing: group g hashes its vertices into its own slice of 2 * (its index count)
// slots (at 2 * index_offsets[g]), keeping the first position of each; a vertex's packed id is
// the number of first uses before its own, so the packed order is the order Sparse gives
// (src/simplifier.cpp:151-199)

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

__global__ void __launch_bounds__(kClodThreads) k_clod_insert(ClodBuffers b)
{
	const uint32_t g = blockIdx.x;
	const uint32_t begin = b.index_offsets[g], end = b.index_offsets[g + 1];
	const uint32_t slots = 2 * (end - begin);
	uint32_t* key = b.slot_key + 2 * (size_t)begin;
	uint32_t* first = b.slot_first + 2 * (size_t)begin;
	for (uint32_t j = begin + threadIdx.x; j < end; j += kClodThreads)
	{
		const uint32_t v = b.indices[j];
		uint32_t slot = clod_slot0(v, slots);
		// at most end - begin distinct keys in twice as many slots: a free or matching slot exists
		// within one sweep of the slice
		for (uint32_t probe = 0; probe < slots; ++probe)
		{
			const uint32_t k = atomicCAS(&key[slot], 0u, v + 1);
			if (k == 0 || k == v + 1)
			{
				atomicMin(&first[slot], j - begin);
				break;
			}
			slot = slot + 1 == slots ? 0 : slot + 1;
		}
	}
}

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
	for (uint32_t w = 0; w < kClodThreads / 64; ++w)
	{
		before += w < wave ? lds[w] : 0;
		all += lds[w];
	}
	__syncthreads();
	total = all;
	return before + incl - v;
}

// vertices per group (its first uses)
__global__ void __launch_bounds__(kClodThreads) k_clod_count(ClodBuffers b)
{
	__shared__ uint32_t lds[kClodThreads / 64];
	const uint32_t g = blockIdx.x;
	const uint32_t begin = b.index_offsets[g], end = b.index_offsets[g + 1];
	const uint32_t slots = 2 * (end - begin);
	const uint32_t* key = b.slot_key + 2 * (size_t)begin;
	const uint32_t* first = b.slot_first + 2 * (size_t)begin;
	uint32_t n = 0;
	for (uint32_t j = begin + threadIdx.x; j < end; j += kClodThreads)
		n += first[clod_find(key, b.indices[j], slots)] == j - begin;
	uint32_t total;
	clod_block_scan(n, lds, total);
	if (threadIdx.x == 0)
		b.unique[g] = total;
}

void launch_clod_hash(hipStream_t s, const ClodBuffers& b)
{
	if (!b.group_count)
		return;
	k_clod_insert<<<b.group_count, kClodThreads, 0, s>>>(b);
	k_clod_count<<<b.group_count, kClodThreads, 0, s>>>(b);
}

// packed ids of the first uses, in order; the packed vertex list, positions and locks
__global__ void __launch_bounds__(kClodThreads) k_clod_first_ids(ClodBuffers b, const float* positions, size_t stride, const uint8_t* locks)
{
	__shared__ uint32_t lds[kClodThreads / 64];
	const uint32_t g = blockIdx.x;
	const uint32_t begin = b.index_offsets[g], end = b.index_offsets[g + 1];
	const uint32_t slots = 2 * (end - begin);
	const uint32_t* key = b.slot_key + 2 * (size_t)begin;
	const uint32_t* first = b.slot_first + 2 * (size_t)begin;
	uint32_t* local = b.slot_local + 2 * (size_t)begin;
	const uint32_t vbase = b.unique[g];
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
			const uint32_t at = vbase + id;
			b.order[at] = v;
			const float* p = positions + (size_t)v * stride;
			b.packed_positions[3 * (size_t)at + 0] = p[0];
			b.packed_positions[3 * (size_t)at + 1] = p[1];
			b.packed_positions[3 * (size_t)at + 2] = p[2];
			b.packed_locks[at] = locks ? locks[v] : 0;
		}
		carry += total;
	}
}

__global__ void __launch_bounds__(kClodThreads) k_clod_local(ClodBuffers b)
{
	const uint32_t g = blockIdx.x;
	const uint32_t begin = b.index_offsets[g], end = b.index_offsets[g + 1];
	const uint32_t slots = 2 * (end - begin);
	const uint32_t* key = b.slot_key + 2 * (size_t)begin;
	const uint32_t* local = b.slot_local + 2 * (size_t)begin;
	for (uint32_t j = begin + threadIdx.x; j < end; j += kClodThreads)
		b.packed_indices[j] = local[clod_find(key, b.indices[j], slots)];
}

void launch_clod_pack(hipStream_t s, const ClodBuffers& b, const float* positions, size_t stride_floats, const uint8_t* locks)
{
	if (!b.group_count)
		return;
	k_clod_first_ids<<<b.group_count, kClodThreads, 0, s>>>(b, positions, stride_floats, locks);
	k_clod_local<<<b.group_count, kClodThreads, 0, s>>>(b);
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

} // namespace mhip
