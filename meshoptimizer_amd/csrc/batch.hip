// Batched simplification: the persistent launch around simplify_patch (batch_device.h).
#include "batch_device.h"

namespace mhip
{

// One workgroup per mesh of one LDS size class; ids lists the class's *count meshes (the grid
// is an upper bound, workgroups past the count leave at once).  At most 80 VGPRs, so that six
// wavefronts share a SIMD and six workgroups a CU when their LDS layouts allow it.  (The largest
// class runs the 1024-thread build of the same device code, batch_wide.hip.)
__global__ void __launch_bounds__(kBatchThreads) __attribute__((amdgpu_waves_per_eu(6))) k_simplify_patches(BatchArgs a, PatchLayout L, const uint32_t* ids, const uint32_t* count)
{
	extern __shared__ __align__(16) uint8_t smem[];
	__shared__ uint32_t part[kBatchThreads / 64];
	__shared__ uint32_t sh[16];
	__shared__ uint64_t box[6];
	__shared__ float shf[8];

	if (blockIdx.x < *count)
		simplify_patch(a, L, ids[blockIdx.x], smem, part, sh, box, shf);
}

// one slot per mesh in its class's id list; the slots of a wavefront's meshes come from one
// atomic per class (same-address atomics from every thread serialise in L2)
__global__ void k_batch_classify(BatchArgs a, BatchClasses c, uint32_t* ids, uint32_t* counts)
{
	size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	int cls = -1;
	if (p < a.patch_count)
	{
		const uint32_t ib = a.index_offsets[p], ie = a.index_offsets[p + 1];
		const uint32_t vb = a.vertex_offsets[p], ve = a.vertex_offsets[p + 1];
		if (ie < ib || ve < vb || ve > a.vertex_count || (ie - ib) % 3 != 0 || a.target_index_counts[p] > ie - ib)
			a.status[p] = 4;
		else
		{
			const uint32_t ni = ie - ib, nv = ve - vb;
			for (int k = 0; k < kBatchClasses && cls < 0; ++k)
				if (nv <= c.vcap[k] && ni <= c.icap[k])
					cls = k;
			if (cls < 0)
				a.status[p] = 1; // too large for the LDS layouts: the plain path
			else if (cls == kBatchClasses - 1)
				a.status[p] = 6; // runs after the other classes (its workgroup takes a whole CU)
		}
	}
	const uint32_t lane = __lane_id();
	const uint64_t below = (1ull << lane) - 1ull;
	for (int k = 0; k < kBatchClasses; ++k)
	{
		const uint64_t mine = __ballot(cls == k);
		if (!mine)
			continue;
		const uint32_t leader = (uint32_t)__builtin_ctzll(mine);
		uint32_t base = 0;
		if (lane == leader)
			base = atomicAdd(&counts[k], (uint32_t)__popcll(mine));
		base = __shfl(base, (int)leader, 64);
		if (cls == k)
			ids[(size_t)k * a.patch_count + base + (uint32_t)__popcll(mine & below)] = (uint32_t)p;
	}
}

// dynamic LDS bytes one workgroup may take so that `per_cu` of them fit a CU (160 KB, allocated
// in 512-byte granules, the kernel's static words included) and the workgroup limit allows it
static uint32_t lds_limit(uint32_t per_cu, uint32_t max_block)
{
	const uint32_t cu_lds = 160 * 1024, granule = 512, static_words = 256;
	uint32_t share = cu_lds / per_cu / granule * granule;
	share = share < max_block ? share : max_block;
	return share - static_words;
}

BatchClasses batch_classes(uint32_t max_block)
{
	// 6, 5, 2 and 1 workgroups per CU (the kernel's 80 VGPRs allow 6 wavefronts per SIMD); the
	// last two classes take meshes with relatively more vertices and hold full candidate arrays
	if (max_block < 64 * 1024)
		max_block = 64 * 1024;
	const uint32_t limits[kBatchClasses] = {lds_limit(6, max_block), lds_limit(5, max_block), lds_limit(2, max_block), lds_limit(1, max_block)};
	const uint32_t vdiv[kBatchClasses] = {4, 4, 3, 3};
	BatchClasses c{};
	for (int k = 0; k < kBatchClasses; ++k)
	{
		if (k && limits[k] <= limits[k - 1])
			continue; // caps 0: no mesh fits
		const bool last = k >= kBatchRetryClass;
		uint32_t vbest = 32, ibest = 96;
		for (uint32_t icap = 96; icap < 0xffffu; icap += 96)
		{
			uint32_t vcap = ((icap / vdiv[k]) + 31) & ~31u;
			uint32_t ccap = last ? icap + 4 : batch_candidate_cap(icap);
			const size_t bytes = k == kBatchClasses - 1 ? batch_wide_lds_bytes(vcap, icap, ccap) : patch_layout(vcap, icap, ccap).total;
			if (bytes > limits[k])
				break;
			vbest = vcap;
			ibest = icap;
		}
		c.vcap[k] = vbest;
		c.icap[k] = ibest;
		c.ccap[k] = last ? ibest + 4 : batch_candidate_cap(ibest);
	}
	return c;
}

void launch_batch_classify(hipStream_t s, const BatchArgs& a, const BatchClasses& classes, uint32_t* ids, uint32_t* counts)
{
	if (a.patch_count)
		k_batch_classify<<<(unsigned int)((a.patch_count + 255) / 256), 256, 0, s>>>(a, classes, ids, counts);
}

#ifdef MESHOPT_BATCH_PROFILE
// the phase counters of both builds of the device code (256 threads here, 1024 in batch_wide.hip),
// summed
extern "C" void meshopt_hip_batch_profile(unsigned long long* out, int reset)
{
	HIP_CHECK(hipMemcpyFromSymbol(out, HIP_SYMBOL(g_batch_profile), sizeof(g_batch_profile)));
	if (reset)
	{
		unsigned long long z[32] = {};
		HIP_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(g_batch_profile), z, sizeof(z)));
	}
	batch_wide_profile_add(out, reset);
}
#endif

// each class's mesh list, larger meshes first (a counting sort by triangle count, one block per
// class): long workgroups start early and the launch ends with short ones
__global__ void __launch_bounds__(1024) k_batch_order(const uint32_t* index_offsets, uint32_t* ids, const uint32_t* counts, size_t P, uint32_t* temp)
{
	__shared__ uint32_t hist[1024];
	__shared__ uint32_t wsum[16];
	const uint32_t tid = threadIdx.x;
	const uint32_t n = counts[blockIdx.x];
	uint32_t* list = ids + (size_t)blockIdx.x * P;
	uint32_t* tmp = temp + (size_t)blockIdx.x * P;
	hist[tid] = 0;
	__syncthreads();
	for (uint32_t i = tid; i < n; i += 1024)
	{
		const uint32_t p = list[i];
		const uint32_t t = (index_offsets[p + 1] - index_offsets[p]) / 3;
		atomicAdd(&hist[1023 - (t < 1023 ? t : 1023)], 1u);
	}
	__syncthreads();
	uint32_t v = hist[tid], incl = v;
	for (int o = 1; o < 64; o <<= 1)
	{
		uint32_t x = __shfl_up(incl, o, 64);
		if (__lane_id() >= (uint32_t)o)
			incl += x;
	}
	if (__lane_id() == 63)
		wsum[tid / 64] = incl;
	__syncthreads();
	uint32_t before = 0;
	for (uint32_t w = 0; w < tid / 64; ++w)
		before += wsum[w];
	__syncthreads();
	hist[tid] = before + incl - v;
	__syncthreads();
	for (uint32_t i = tid; i < n; i += 1024)
	{
		const uint32_t p = list[i];
		const uint32_t t = (index_offsets[p + 1] - index_offsets[p]) / 3;
		tmp[atomicAdd(&hist[1023 - (t < 1023 ? t : 1023)], 1u)] = p;
	}
	__syncthreads();
	for (uint32_t i = tid; i < n; i += 1024)
		list[i] = tmp[i];
}

void launch_batch_order(hipStream_t s, const BatchArgs& a, uint32_t* ids, const uint32_t* counts, uint32_t* temp)
{
	if (a.patch_count)
		k_batch_order<<<kBatchClasses, 1024, 0, s>>>(a.index_offsets, ids, counts, a.patch_count, temp);
}

bool batch_allow_lds(size_t bytes)
{
	return hipFuncSetAttribute(reinterpret_cast<const void*>(&k_simplify_patches), hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) ==
	       hipSuccess;
}

size_t batch_lds_bytes(uint32_t vmax, uint32_t imax, uint32_t cmax)
{
	return patch_layout(vmax, imax, cmax).total;
}

void launch_batch(hipStream_t s, const BatchArgs& a, uint32_t vmax, uint32_t imax, uint32_t cmax, const uint32_t* ids, const uint32_t* count_dev,
                  size_t grid)
{
	PatchLayout L = patch_layout(vmax, imax, cmax);
	if (grid)
		k_simplify_patches<<<(unsigned int)grid, kBatchThreads, L.total, s>>>(a, L, ids, count_dev);
}

} // namespace mhip
