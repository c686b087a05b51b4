// The launch forms of the batched kernel for the workgroup width one translation unit builds
// (MHIP_BATCH_THREADS, see batch_device.h), and their host helpers.  Each width's symbols live in
// that width's inline namespace; batch.hip dispatches on the width a size class asks for.
#pragma once

#include "batch_device.h"
#include "prims.h"

// from 128 to 512 threads the kernel is held to 80 VGPRs, so that six wavefronts share a SIMD; a
// 1024-thread workgroup has four per SIMD and may take 128
#if MHIP_BATCH_THREADS < 1024
#define MHIP_BATCH_OCCUPANCY __attribute__((amdgpu_waves_per_eu(6)))
#else
#define MHIP_BATCH_OCCUPANCY
#endif

namespace mhip
{
inline namespace MHIP_BATCH_NS(MHIP_BATCH_THREADS)
{

// one workgroup per mesh of one LDS size class: ids lists the class's *count meshes (the grid is
// an upper bound, workgroups past the count leave at once).  kAttr: the build for batches with
// attributes
template <bool kAttr>
__global__ void __launch_bounds__(kBatchThreads) MHIP_BATCH_OCCUPANCY k_simplify_patches(BatchArgs a, PatchLayout L, const uint32_t* ids, const uint32_t* count)
{
	extern __shared__ __align__(16) uint8_t smem[];
	__shared__ uint32_t part[kBatchWaves];
	__shared__ uint32_t sh[16];
	__shared__ uint64_t box[6];
	__shared__ float shf[8];

	if (blockIdx.x < *count)
		simplify_patch<kAttr>(a, L, ids[blockIdx.x], smem, part, sh, box, shf);
}

// each workgroup takes the next mesh of the list from *cursor (zero at the launch) until the list
// ends, so a workgroup that finishes early takes the next largest mesh (simplify_from_cursor)
template <bool kAttr>
__global__ void __launch_bounds__(kBatchThreads) MHIP_BATCH_OCCUPANCY k_simplify_patches_cursor(BatchArgs a, PatchLayout L, const uint32_t* ids,
                                                                                               const uint32_t* count, uint32_t* cursor)
{
	extern __shared__ __align__(16) uint8_t smem[];
	__shared__ uint32_t part[kBatchWaves];
	__shared__ uint32_t sh[16];
	__shared__ uint64_t box[6];
	__shared__ float shf[8];
	__shared__ uint32_t next;

	simplify_from_cursor<kAttr>(a, L, ids, *count, cursor, smem, part, sh, box, shf, &next);
}

size_t width_lds_bytes(uint32_t vmax, uint32_t imax, uint32_t cmax)
{
	return patch_layout(vmax, imax, cmax).total;
}

static bool allow_lds(const void* kernel, size_t bytes)
{
	return hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) == hipSuccess;
}

// every form may take `bytes` of dynamic LDS; false when the runtime refuses
bool width_allow_lds(size_t bytes)
{
	return allow_lds(reinterpret_cast<const void*>(&k_simplify_patches<false>), bytes) &&
	       allow_lds(reinterpret_cast<const void*>(&k_simplify_patches_cursor<false>), bytes) &&
	       allow_lds(reinterpret_cast<const void*>(&k_simplify_patches<true>), bytes) &&
	       allow_lds(reinterpret_cast<const void*>(&k_simplify_patches_cursor<true>), bytes);
}

// cursor == null: one workgroup per mesh slot (grid at least the class's count); otherwise `grid`
// workgroups loop over the list through the cursor word, cleared here on the stream
void width_launch(hipStream_t s, const BatchArgs& a, uint32_t vmax, uint32_t imax, uint32_t cmax, const uint32_t* ids, const uint32_t* count_dev,
                  size_t grid, uint32_t* cursor)
{
	const PatchLayout L = patch_layout(vmax, imax, cmax);
	if (!grid)
		return;
	const bool attr = a.attribute_count != 0;
	if (cursor)
	{
		fill_u32(s, cursor, 0u, 1);
		if (attr)
			k_simplify_patches_cursor<true><<<(unsigned int)grid, kBatchThreads, L.total, s>>>(a, L, ids, count_dev, cursor);
		else
			k_simplify_patches_cursor<false><<<(unsigned int)grid, kBatchThreads, L.total, s>>>(a, L, ids, count_dev, cursor);
	}
	else if (attr)
		k_simplify_patches<true><<<(unsigned int)grid, kBatchThreads, L.total, s>>>(a, L, ids, count_dev);
	else
		k_simplify_patches<false><<<(unsigned int)grid, kBatchThreads, L.total, s>>>(a, L, ids, count_dev);
}

#ifdef MESHOPT_BATCH_PROFILE
// adds this width's phase counters (MESHOPT_BATCH_PROFILE builds) to out[32]
void width_profile_add(unsigned long long* out, int reset)
{
	unsigned long long w[32];
	HIP_CHECK(hipMemcpyFromSymbol(w, HIP_SYMBOL(g_batch_profile), sizeof(w)));
	for (int k = 0; k < 32; ++k)
		out[k] += w[k];
	if (reset)
	{
		unsigned long long z[32] = {};
		HIP_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(g_batch_profile), z, sizeof(z)));
	}
}
#endif

} // inline namespace
} // namespace mhip
