// Exclusive scan and stable LSD radix sort for gfx950 (wave64, 256-thread workgroups).
#include "prims.h"

namespace mhip
{

__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t v)
{
	const int lane = __lane_id();
#pragma unroll
	for (int o = 1; o < 64; o <<= 1)
	{
		uint32_t t = __shfl_up(v, o, 64);
		if (lane >= o)
			v += t;
	}
	return v;
}

// exclusive scan across a 256-thread block; lds must hold 4 words
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* lds, uint32_t& total)
{
	uint32_t incl = wave_inclusive_scan(v);
	const int wave = threadIdx.x >> 6;
	if (__lane_id() == 63)
		lds[wave] = incl;
	__syncthreads();
	uint32_t before = 0, all = 0;
#pragma unroll
	for (int w = 0; w < kScanThreads / 64; ++w)
	{
		uint32_t c = lds[w];
		before += (w < wave) ? c : 0;
		all += c;
	}
	__syncthreads();
	total = all;
	return before + incl - v;
}

__device__ __forceinline__ uint32_t block_reduce_sum(uint32_t v, uint32_t* lds)
{
	uint32_t total;
	block_exclusive_scan(v, lds, total);
	return total;
}

__device__ __forceinline__ size_t live_count(const uint32_t* n_dev, size_t n_max)
{
	if (!n_dev)
		return n_max;
	size_t n = *n_dev;
	return n < n_max ? n : n_max;
}

__global__ void __launch_bounds__(kScanThreads) k_scan_reduce(const uint32_t* in, size_t n_max, const uint32_t* n_dev, uint32_t* block_sums)
{
	__shared__ uint32_t lds[4];
	size_t n = live_count(n_dev, n_max);
	size_t base = (size_t)blockIdx.x * kScanTile;
	uint32_t sum = 0;
#pragma unroll 4
	for (int j = 0; j < kScanItems; ++j)
	{
		size_t e = base + (size_t)j * kScanThreads + threadIdx.x;
		sum += e < n ? in[e] : 0u;
	}
	sum = block_reduce_sum(sum, lds);
	if (threadIdx.x == 0)
		block_sums[blockIdx.x] = sum;
}

// single block: exclusive scan of the block sums in place, total to *total
__global__ void __launch_bounds__(kScanThreads) k_scan_blocks(uint32_t* sums, uint32_t count, uint32_t* total)
{
	__shared__ uint32_t lds[4];
	uint32_t carry = 0;
	for (uint32_t base = 0; base < count; base += kScanThreads)
	{
		uint32_t e = base + threadIdx.x;
		uint32_t v = e < count ? sums[e] : 0u;
		uint32_t tot;
		uint32_t ex = block_exclusive_scan(v, lds, tot);
		if (e < count)
			sums[e] = carry + ex;
		carry += tot;
	}
	if (threadIdx.x == 0 && total)
		*total = carry;
}

__global__ void __launch_bounds__(kScanThreads) k_scan_apply(const uint32_t* in, uint32_t* out, size_t n_max, const uint32_t* n_dev, const uint32_t* block_offsets)
{
	__shared__ uint32_t lds[4];
	size_t n = live_count(n_dev, n_max);
	size_t base = (size_t)blockIdx.x * kScanTile;
	if (base >= n)
		return;
	uint32_t carry = block_offsets[blockIdx.x];
	for (int j = 0; j < kScanItems; ++j)
	{
		size_t e = base + (size_t)j * kScanThreads + threadIdx.x;
		uint32_t v = e < n ? in[e] : 0u;
		uint32_t tot;
		uint32_t ex = block_exclusive_scan(v, lds, tot);
		if (e < n)
			out[e] = carry + ex;
		carry += tot;
	}
}

size_t scan_temp_words(size_t n)
{
	return div_up(n ? n : 1, kScanTile) + 16;
}

void scan_exclusive(hipStream_t s, const uint32_t* in, uint32_t* out, size_t n, const uint32_t* n_dev, uint32_t* total, uint32_t* temp)
{
	unsigned int blocks = div_up(n ? n : 1, kScanTile);
	k_scan_reduce<<<blocks, kScanThreads, 0, s>>>(in, n, n_dev, temp);
	k_scan_blocks<<<1, kScanThreads, 0, s>>>(temp, blocks, total);
	k_scan_apply<<<blocks, kScanThreads, 0, s>>>(in, out, n, n_dev, temp);
}

// ------------------------------------------------------------------------------------------
// stable LSD radix sort, 8-bit digits

__global__ void __launch_bounds__(kSortThreads) k_radix_hist(const uint32_t* keys, size_t n_max, const uint32_t* n_dev, int shift, uint32_t* counts)
{
	__shared__ uint32_t hist[256];
	size_t n = live_count(n_dev, n_max);
	hist[threadIdx.x] = 0;
	__syncthreads();
	size_t base = (size_t)blockIdx.x * kSortTile;
#pragma unroll 4
	for (int j = 0; j < kSortItems; ++j)
	{
		size_t e = base + (size_t)j * kSortThreads + threadIdx.x;
		if (e < n)
			atomicAdd(&hist[(keys[e] >> shift) & 255u], 1u);
	}
	__syncthreads();
	counts[(size_t)threadIdx.x * gridDim.x + blockIdx.x] = hist[threadIdx.x];
}

__global__ void __launch_bounds__(kSortThreads) k_radix_scatter(const uint32_t* keys, const uint32_t* vals, uint32_t* keys_out, uint32_t* vals_out, size_t n_max, const uint32_t* n_dev, int shift, const uint32_t* offsets)
{
	__shared__ uint32_t run[256];
	__shared__ uint32_t woff[2][4][256];

	size_t n = live_count(n_dev, n_max);
	size_t base = (size_t)blockIdx.x * kSortTile;
	if (base >= n)
		return;

	const int t = threadIdx.x;
	const int wave = t >> 6;
	const int lane = __lane_id();

	run[t] = offsets[(size_t)t * gridDim.x + blockIdx.x];
#pragma unroll
	for (int w = 0; w < 4; ++w)
	{
		woff[0][w][t] = 0;
		woff[1][w][t] = 0;
	}
	__syncthreads();

	for (int j = 0; j < kSortItems; ++j)
	{
		uint32_t* off = &woff[j & 1][0][0];
		uint32_t* off_next = &woff[(j + 1) & 1][0][0];

		size_t e = base + (size_t)j * kSortThreads + t;
		bool valid = e < n;
		uint32_t key = valid ? keys[e] : 0u;
		uint32_t d = (key >> shift) & 255u;

		uint64_t peers = __ballot(valid);
#pragma unroll
		for (int b = 0; b < 8; ++b)
		{
			uint64_t bb = __ballot((d >> b) & 1u);
			peers &= ((d >> b) & 1u) ? bb : ~bb;
		}
		uint64_t lower = peers & ((1ull << lane) - 1ull);
		uint32_t rank = (uint32_t)__popcll(lower);
		if (valid && (63 - __clzll((long long)peers)) == lane)
			off[wave * 256 + d] = (uint32_t)__popcll(peers);
		__syncthreads();

		{
			uint32_t c0 = off[0 * 256 + t], c1 = off[1 * 256 + t], c2 = off[2 * 256 + t], c3 = off[3 * 256 + t];
			uint32_t r = run[t];
			off[0 * 256 + t] = r;
			off[1 * 256 + t] = r + c0;
			off[2 * 256 + t] = r + c0 + c1;
			off[3 * 256 + t] = r + c0 + c1 + c2;
			run[t] = r + c0 + c1 + c2 + c3;
#pragma unroll
			for (int w = 0; w < 4; ++w)
				off_next[w * 256 + t] = 0;
		}
		__syncthreads();

		if (valid)
		{
			uint32_t pos = off[wave * 256 + d] + rank;
			keys_out[pos] = key;
			vals_out[pos] = vals[e];
		}
	}
}

size_t sort_temp_words(size_t n)
{
	size_t blocks = div_up(n ? n : 1, kSortTile);
	return 256 * blocks + scan_temp_words(256 * blocks) + 16;
}

void radix_sort_pairs(hipStream_t s, uint32_t* keys, uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt, size_t n, const uint32_t* n_dev, int bits, uint32_t* temp)
{
	if (n == 0 || bits <= 0)
		return;
	unsigned int blocks = div_up(n, kSortTile);
	uint32_t* counts = temp;
	uint32_t* scan_tmp = temp + 256 * (size_t)blocks;

	uint32_t *ka = keys, *va = vals, *kb = keys_alt, *vb = vals_alt;
	int passes = (bits + 7) / 8;
	for (int p = 0; p < passes; ++p)
	{
		int shift = p * 8;
		k_radix_hist<<<blocks, kSortThreads, 0, s>>>(ka, n, n_dev, shift, counts);
		scan_exclusive(s, counts, counts, 256 * (size_t)blocks, nullptr, nullptr, scan_tmp);
		k_radix_scatter<<<blocks, kSortThreads, 0, s>>>(ka, va, kb, vb, n, n_dev, shift, counts);
		uint32_t* t;
		t = ka, ka = kb, kb = t;
		t = va, va = vb, vb = t;
	}
	if (ka != keys)
	{
		// odd pass count: the result sits in the alt buffers
		size_t bytes = n * sizeof(uint32_t);
		HIP_CHECK(hipMemcpyAsync(keys, ka, bytes, hipMemcpyDeviceToDevice, s));
		HIP_CHECK(hipMemcpyAsync(vals, va, bytes, hipMemcpyDeviceToDevice, s));
	}
}

__global__ void k_fill_u32(uint32_t* p, uint32_t value, size_t n)
{
	size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		p[i] = value;
}

__global__ void k_iota_u32(uint32_t* p, size_t n)
{
	size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		p[i] = (uint32_t)i;
}

void fill_u32(hipStream_t s, uint32_t* p, uint32_t value, size_t n)
{
	if (n)
		k_fill_u32<<<div_up(n, 256), 256, 0, s>>>(p, value, n);
}

void iota_u32(hipStream_t s, uint32_t* p, size_t n)
{
	if (n)
		k_iota_u32<<<div_up(n, 256), 256, 0, s>>>(p, n);
}

} // namespace mhip
