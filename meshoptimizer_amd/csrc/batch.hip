// Batched simplification: the persistent launch around simplify_patch (batch_device.h).
#include "batch_device.h"

namespace mhip
{

// One workgroup per mesh of one LDS size class; ids lists the class's meshes.
__global__ void __launch_bounds__(kBatchThreads) k_simplify_patches(BatchArgs a, PatchLayout L, const uint32_t* ids)
{
	extern __shared__ __align__(16) uint8_t smem[];
	__shared__ uint32_t part[kBatchThreads / 64];
	__shared__ uint32_t sh[16];
	__shared__ uint64_t box[6];
	__shared__ float shf[8];

	const uint32_t patch = ids[blockIdx.x];
	simplify_patch(a, L, patch, smem, part, sh, box, shf);
}

#ifdef MESHOPT_BATCH_PROFILE
extern "C" void meshopt_hip_batch_profile(unsigned long long* out, int reset)
{
	HIP_CHECK(hipMemcpyFromSymbol(out, HIP_SYMBOL(g_batch_profile), sizeof(g_batch_profile)));
	if (reset)
	{
		unsigned long long z[32] = {};
		HIP_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(g_batch_profile), z, sizeof(z)));
	}
}
#endif

size_t batch_lds_bytes(uint32_t vmax, uint32_t imax)
{
	return patch_layout(vmax, imax).total;
}

void launch_batch(hipStream_t s, const BatchArgs& a, uint32_t vmax, uint32_t imax, const uint32_t* ids, size_t count)
{
	PatchLayout L = patch_layout(vmax, imax);
	if (count)
		k_simplify_patches<<<(unsigned int)count, kBatchThreads, L.total, s>>>(a, L, ids);
}

} // namespace mhip
