// Batched simplification: the size classes, their launch at each workgroup width, and the
// classify and order kernels.  This translation unit builds the 256-thread device code
// (batch_device.h); batch_narrow.hip, batch_mid.hip and batch_wide.hip build 128, 512 and 1024.
#define MHIP_BATCH_THREADS 256
#include "batch_kernels.h"

namespace mhip
{

// the other widths' helpers (batch_kernels.h, built in batch_narrow.hip, batch_mid.hip and
// batch_wide.hip)
namespace threads128
{
size_t width_lds_bytes(uint32_t vmax, uint32_t imax, uint32_t cmax);
bool width_allow_lds(size_t bytes);
void width_launch(hipStream_t s, const BatchArgs& a, uint32_t vmax, uint32_t imax, uint32_t cmax, const uint32_t* ids, const uint32_t* count_dev,
                  size_t grid, uint32_t* cursor);
#ifdef MESHOPT_BATCH_PROFILE
void width_profile_add(unsigned long long* out, int reset);
#endif
} // namespace threads128
namespace threads512
{
size_t width_lds_bytes(uint32_t vmax, uint32_t imax, uint32_t cmax);
bool width_allow_lds(size_t bytes);
void width_launch(hipStream_t s, const BatchArgs& a, uint32_t vmax, uint32_t imax, uint32_t cmax, const uint32_t* ids, const uint32_t* count_dev,
                  size_t grid, uint32_t* cursor);
#ifdef MESHOPT_BATCH_PROFILE
void width_profile_add(unsigned long long* out, int reset);
#endif
} // namespace threads512
namespace threads1024
{
size_t width_lds_bytes(uint32_t vmax, uint32_t imax, uint32_t cmax);
bool width_allow_lds(size_t bytes);
void width_launch(hipStream_t s, const BatchArgs& a, uint32_t vmax, uint32_t imax, uint32_t cmax, const uint32_t* ids, const uint32_t* count_dev,
                  size_t grid, uint32_t* cursor);
#ifdef MESHOPT_BATCH_PROFILE
void width_profile_add(unsigned long long* out, int reset);
#endif
} // namespace threads1024

size_t batch_lds_bytes(uint32_t threads, uint32_t vmax, uint32_t imax, uint32_t cmax)
{
	if (threads == 1024)
		return threads1024::width_lds_bytes(vmax, imax, cmax);
	if (threads == 512)
		return threads512::width_lds_bytes(vmax, imax, cmax);
	if (threads == 128)
		return threads128::width_lds_bytes(vmax, imax, cmax);
	return threads256::width_lds_bytes(vmax, imax, cmax);
}

bool batch_allow_lds(uint32_t threads, size_t bytes)
{
	if (threads == 1024)
		return threads1024::width_allow_lds(bytes);
	if (threads == 512)
		return threads512::width_allow_lds(bytes);
	if (threads == 128)
		return threads128::width_allow_lds(bytes);
	return threads256::width_allow_lds(bytes);
}

void launch_batch(hipStream_t s, uint32_t threads, const BatchArgs& a, uint32_t vmax, uint32_t imax, uint32_t cmax, const uint32_t* ids,
                  const uint32_t* count_dev, size_t grid, uint32_t* cursor)
{
	if (threads == 1024)
		threads1024::width_launch(s, a, vmax, imax, cmax, ids, count_dev, grid, cursor);
	else if (threads == 512)
		threads512::width_launch(s, a, vmax, imax, cmax, ids, count_dev, grid, cursor);
	else if (threads == 128)
		threads128::width_launch(s, a, vmax, imax, cmax, ids, count_dev, grid, cursor);
	else
		threads256::width_launch(s, a, vmax, imax, cmax, ids, count_dev, grid, cursor);
}

// one slot per mesh in its class's id list; the slots of a wavefront's meshes come from one
// atomic per class (same-address atomics from every thread serialise in L2)
// the size class of mesh p (-1: none) and the status word to write for it (0: none; the kernel
// of its class writes it), with its triangle count
__device__ __forceinline__ int batch_mesh_class(const BatchArgs& a, const BatchClasses& c, size_t p, uint32_t& status, uint32_t& tris)
{
	const uint32_t ib = a.index_offsets[p], ie = a.index_offsets[p + 1];
	const uint32_t vb = a.vertex_offsets[p], ve = a.vertex_offsets[p + 1];
	status = 0;
	tris = 0;
	if (ie < ib || ve < vb || ve > a.vertex_count || (ie - ib) % 3 != 0 || a.target_index_counts[p] > ie - ib)
	{
		status = 4;
		return -1;
	}
	const uint32_t ni = ie - ib, nv = ve - vb;
	tris = ni / 3;
	int cls = -1;
	for (int k = 0; k < kBatchClasses && cls < 0; ++k)
		if (nv <= c.vcap[k] && ni <= c.icap[k])
			cls = k;
	if (cls < 0)
		status = 1; // too large for the LDS layouts: the plain path
	else if (cls == kBatchClasses - 1)
		status = 6; // runs after the other classes (its workgroup takes a whole CU)
	return cls;
}

__global__ void k_batch_classify(BatchArgs a, BatchClasses c, uint32_t* ids, uint32_t* counts)
{
	size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	int cls = -1;
	if (p < a.patch_count)
	{
		uint32_t status, tris;
		cls = batch_mesh_class(a, c, p, status, tris);
		if (status)
			a.status[p] = status;
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

BatchShape batch_default_shape()
{
	// 6, 5, 2 and 1 workgroups per CU: the 80-VGPR 256-thread kernel allows 6 wavefronts per SIMD;
	// the last class takes a whole CU with 1024 threads
	BatchShape s{};
	const uint32_t per_cu[kBatchClasses] = {6, 5, 2, 1}, threads[kBatchClasses] = {256, 256, 256, 1024};
	for (int k = 0; k < kBatchClasses; ++k)
	{
		s.per_cu[k] = per_cu[k];
		s.threads[k] = threads[k];
		s.cursor[k] = k == kBatchClasses - 1;
	}
	return s;
}

const char* batch_shape_error(const BatchShape& s)
{
	for (int k = 0; k < kBatchClasses; ++k)
	{
		const uint32_t t = s.threads[k];
		if (t != 128 && t != 256 && t != 512 && t != 1024)
			return "threads per workgroup must be 128, 256, 512 or 1024";
		if (s.per_cu[k] < 1)
			return "every class needs at least one workgroup per CU";
		// wavefronts per SIMD: 6 at most for the 80-VGPR widths, 4 for the 1024-thread one
		const uint32_t per_simd = (s.per_cu[k] * (t / 64) + 3) / 4;
		if (per_simd > (t == 1024 ? 4u : 6u))
			return "more wavefronts per SIMD than the width's register budget allows";
		if (k && s.per_cu[k] > s.per_cu[k - 1])
			return "workgroups per CU must not grow from one class to the next";
	}
	if (s.per_cu[kBatchClasses - 1] != 1)
		return "the last class takes a whole CU (one workgroup per CU)";
	return nullptr;
}

BatchClasses batch_classes(uint32_t max_block, const BatchShape& shape)
{
	if (max_block < 64 * 1024)
		max_block = 64 * 1024;
	const uint32_t vdiv[kBatchClasses] = {4, 4, 4, 3};
	BatchClasses c{};
	c.shape = shape;
	uint32_t prev_icap = 0;
	for (int k = 0; k < kBatchClasses; ++k)
	{
		const uint32_t limit = lds_limit(shape.per_cu[k], max_block);
		// classes from the retry class on hold the full candidate bound (boundEdgeCollapses)
		const bool last = k >= kBatchRetryClass;
		uint32_t vbest = 0, ibest = 0;
		for (uint32_t icap = 96; icap < 0xffffu; icap += 96)
		{
			uint32_t vcap = ((icap / vdiv[k]) + 31) & ~31u;
			uint32_t ccap = last ? icap + 4 : batch_candidate_cap(icap);
			if (batch_lds_bytes(shape.threads[k], vcap, icap, ccap) > limit)
				break;
			vbest = vcap;
			ibest = icap;
		}
		if (ibest <= prev_icap)
			continue; // caps 0: no mesh fits
		c.vcap[k] = vbest;
		c.icap[k] = ibest;
		c.ccap[k] = last ? ibest + 4 : batch_candidate_cap(ibest);
		prev_icap = ibest;
	}
	return c;
}

void launch_batch_classify(hipStream_t s, const BatchArgs& a, const BatchClasses& classes, uint32_t* ids, uint32_t* counts)
{
	if (a.patch_count)
		k_batch_classify<<<(unsigned int)((a.patch_count + 255) / 256), 256, 0, s>>>(a, classes, ids, counts);
}

#ifdef MESHOPT_BATCH_PROFILE
// the phase counters of the four builds of the device code (128, 256, 512 and 1024 threads), summed
extern "C" void meshopt_hip_batch_profile(unsigned long long* out, int reset)
{
	for (int k = 0; k < 32; ++k)
		out[k] = 0;
	threads128::width_profile_add(out, reset);
	threads256::width_profile_add(out, reset);
	threads512::width_profile_add(out, reset);
	threads1024::width_profile_add(out, reset);
}
#endif

// exclusive scan of hist[1024] in place by a 1024-thread block; returns the total
__device__ __forceinline__ uint32_t order_scan(uint32_t* hist, uint32_t* wsum)
{
	const uint32_t tid = threadIdx.x;
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
	uint32_t before = 0, total = 0;
	for (uint32_t w = 0; w < 16; ++w)
	{
		before += w < tid / 64 ? wsum[w] : 0u;
		total += wsum[w];
	}
	__syncthreads();
	hist[tid] = before + incl - v;
	__syncthreads();
	return total;
}

// classification and ordering in one launch, for batches of up to kFusedOrderMax meshes:
// workgroup k keeps the meshes of class k (every workgroup classifies every mesh, the first
// writes the status words) and lists them larger first with a counting sort by triangle count
__global__ void __launch_bounds__(1024) k_batch_classify_order(BatchArgs a, BatchClasses c, uint32_t* ids, uint32_t* counts, uint32_t* host_counts)
{
	__shared__ uint32_t hist[1024];
	__shared__ uint32_t wsum[16];
	// each mesh's histogram bucket when it is of this workgroup's class, else kNone16: the
	// placement pass reads it here instead of classifying again
	__shared__ uint16_t bucket[kFusedOrderMax];
	const uint32_t tid = threadIdx.x, k = blockIdx.x;
	const size_t P = a.patch_count;
	uint32_t* list = ids + (size_t)k * P;
	hist[tid] = 0;
	__syncthreads();
	for (size_t p = tid; p < P; p += 1024)
	{
		uint32_t status, tris;
		const int cls = batch_mesh_class(a, c, p, status, tris);
		if (k == 0 && status)
			a.status[p] = status;
		uint16_t b = kNone16;
		if (cls == (int)k)
		{
			b = (uint16_t)(1023 - (tris < 1023 ? tris : 1023));
			atomicAdd(&hist[b], 1u);
		}
		bucket[p] = b;
	}
	__syncthreads();
	const uint32_t total = order_scan(hist, wsum);
	for (size_t p = tid; p < P; p += 1024)
	{
		const uint16_t b = bucket[p];
		if (b != kNone16)
			list[atomicAdd(&hist[b], 1u)] = (uint32_t)p;
	}
	if (tid == 0)
	{
		counts[k] = total;
		if (host_counts)
			host_counts[k] = total;
	}
}

bool launch_batch_classify_order(hipStream_t s, const BatchArgs& a, const BatchClasses& classes, uint32_t* ids, uint32_t* counts, uint32_t* host_counts)
{
	if (a.patch_count > kFusedOrderMax)
		return false;
	if (a.patch_count)
		k_batch_classify_order<<<kBatchClasses, 1024, 0, s>>>(a, classes, ids, counts, host_counts);
	else
		fill_u32(s, counts, 0u, kBatchClasses);
	return true;
}

// each class's mesh list, larger meshes first (a counting sort by triangle count, one block per
// class): long workgroups start early and the launch ends with short ones
__global__ void __launch_bounds__(1024) k_batch_order(const uint32_t* index_offsets, uint32_t* ids, const uint32_t* counts, size_t P, uint32_t* temp,
                                                      uint32_t* host_counts)
{
	__shared__ uint32_t hist[1024];
	__shared__ uint32_t wsum[16];
	const uint32_t tid = threadIdx.x;
	const uint32_t n = counts[blockIdx.x];
	if (host_counts && tid == 0)
		host_counts[blockIdx.x] = n;
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
	order_scan(hist, wsum);
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

void launch_batch_order(hipStream_t s, const BatchArgs& a, uint32_t* ids, const uint32_t* counts, uint32_t* temp, uint32_t* host_counts)
{
	if (a.patch_count)
		k_batch_order<<<kBatchClasses, 1024, 0, s>>>(a.index_offsets, ids, counts, a.patch_count, temp, host_counts);
}

} // namespace mhip
