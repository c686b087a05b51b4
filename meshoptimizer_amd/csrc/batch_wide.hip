// The batched kernel's device code built for 1024-thread workgroups (16 wavefronts), for the size
// class whose workgroup takes a whole CU's LDS: with 256 threads each SIMD would hold a single
// wavefront.  The device code is the same source as batch.hip's, in its own inline namespace.
#define MHIP_BATCH_THREADS 1024
#include "batch_device.h"
#include "prims.h"

namespace mhip
{

// each workgroup takes the next mesh of the list from *cursor (zero at the launch) until the
// list ends, so a workgroup that finishes early takes the next largest mesh (simplify_from_cursor)
__global__ void __launch_bounds__(kBatchThreads) k_simplify_patches_wide(BatchArgs a, PatchLayout L, const uint32_t* ids, const uint32_t* count,
                                                                         uint32_t* cursor)
{
	extern __shared__ __align__(16) uint8_t smem[];
	__shared__ uint32_t part[kBatchWaves];
	__shared__ uint32_t sh[16];
	__shared__ uint64_t box[6];
	__shared__ float shf[8];
	__shared__ uint32_t next;

	simplify_from_cursor(a, L, ids, *count, cursor, smem, part, sh, box, shf, &next);
}

#ifdef MESHOPT_BATCH_PROFILE
void batch_wide_profile_add(unsigned long long* out, int reset)
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

size_t batch_wide_lds_bytes(uint32_t vmax, uint32_t imax, uint32_t cmax)
{
	return patch_layout(vmax, imax, cmax).total;
}

bool batch_wide_allow_lds(size_t bytes)
{
	return hipFuncSetAttribute(reinterpret_cast<const void*>(&k_simplify_patches_wide), hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) ==
	       hipSuccess;
}

void launch_batch_wide(hipStream_t s, const BatchArgs& a, uint32_t vmax, uint32_t imax, uint32_t cmax, const uint32_t* ids, const uint32_t* count_dev,
                       size_t grid, uint32_t* cursor)
{
	PatchLayout L = patch_layout(vmax, imax, cmax);
	if (!grid)
		return;
	fill_u32(s, cursor, 0u, 1);
	k_simplify_patches_wide<<<(unsigned int)grid, kBatchThreads, L.total, s>>>(a, L, ids, count_dev, cursor);
}

} // namespace mhip
