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

// Tiles of kScanTile elements, one per block; thread t owns the kScanItems consecutive elements
// at tile offset t * kScanItems (16-byte vector accesses when the arrays allow them, scalar ones
// for the tail), so a tile needs one block scan instead of one per item.
template <bool Vec>
__device__ __forceinline__ void load_run(const uint32_t* in, size_t e, size_t n, uint32_t v[kScanItems])
{
	if (Vec && e + kScanItems <= n)
	{
		const uint4* p = reinterpret_cast<const uint4*>(in + e);
#pragma unroll
		for (int q = 0; q < kScanItems / 4; ++q)
		{
			uint4 x = p[q];
			v[4 * q + 0] = x.x;
			v[4 * q + 1] = x.y;
			v[4 * q + 2] = x.z;
			v[4 * q + 3] = x.w;
		}
	}
	else
	{
#pragma unroll
		for (int k = 0; k < kScanItems; ++k)
			v[k] = e + k < n ? in[e + k] : 0u;
	}
}

template <bool Vec>
__global__ void __launch_bounds__(kScanThreads) k_scan_reduce(const uint32_t* in, size_t n_max, const uint32_t* n_dev, uint32_t* block_sums)
{
	__shared__ uint32_t lds[4];
	const size_t n = live_count(n_dev, n_max);
	const size_t e = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanItems;
	uint32_t v[kScanItems];
	load_run<Vec>(in, e, n, v);
	uint32_t sum = 0;
#pragma unroll
	for (int k = 0; k < kScanItems; ++k)
		sum += v[k];
	sum = block_reduce_sum(sum, lds);
	if (threadIdx.x == 0)
		block_sums[blockIdx.x] = sum;
}

// each block adds the sums of the tiles before it (read directly: a few per thread), scans its
// tile and writes it; the last block also writes the total
template <bool Vec>
__global__ void __launch_bounds__(kScanThreads) k_scan_apply(const uint32_t* in, uint32_t* out, size_t n_max, const uint32_t* n_dev, const uint32_t* block_sums,
                                                           uint32_t* total)
{
	__shared__ uint32_t lds[4];
	const size_t n = live_count(n_dev, n_max);
	uint32_t before = 0;
	for (uint32_t b = threadIdx.x; b < blockIdx.x; b += kScanThreads)
		before += block_sums[b];
	before = block_reduce_sum(before, lds);
	if (blockIdx.x + 1 == gridDim.x && threadIdx.x == 0 && total)
		*total = before + block_sums[blockIdx.x];

	const size_t e = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanItems;
	if ((size_t)blockIdx.x * kScanTile >= n)
		return; // uniform per block: no barrier follows for the blocks that leave
	uint32_t v[kScanItems];
	load_run<Vec>(in, e, n, v);
	uint32_t run = 0;
#pragma unroll
	for (int k = 0; k < kScanItems; ++k)
		run += v[k];
	uint32_t tile_total;
	uint32_t carry = before + block_exclusive_scan(run, lds, tile_total);
	if (Vec && e + kScanItems <= n)
	{
		uint4* p = reinterpret_cast<uint4*>(out + e);
#pragma unroll
		for (int q = 0; q < kScanItems / 4; ++q)
		{
			uint4 x;
			x.x = carry;
			carry += v[4 * q + 0];
			x.y = carry;
			carry += v[4 * q + 1];
			x.z = carry;
			carry += v[4 * q + 2];
			x.w = carry;
			carry += v[4 * q + 3];
			p[q] = x;
		}
	}
	else
	{
#pragma unroll
		for (int k = 0; k < kScanItems; ++k)
		{
			if (e + k < n)
				out[e + k] = carry;
			carry += v[k];
		}
	}
}

size_t scan_temp_words(size_t n)
{
	return div_up(n ? n : 1, kScanTile) + 16;
}

void scan_exclusive(hipStream_t s, const uint32_t* in, uint32_t* out, size_t n, const uint32_t* n_dev, uint32_t* total, uint32_t* temp)
{
	unsigned int blocks = div_up(n ? n : 1, kScanTile);
	const bool vec = ((reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(out)) & 15) == 0;
	if (vec)
	{
		k_scan_reduce<true><<<blocks, kScanThreads, 0, s>>>(in, n, n_dev, temp);
		k_scan_apply<true><<<blocks, kScanThreads, 0, s>>>(in, out, n, n_dev, temp, total);
	}
	else
	{
		k_scan_reduce<false><<<blocks, kScanThreads, 0, s>>>(in, n, n_dev, temp);
		k_scan_apply<false><<<blocks, kScanThreads, 0, s>>>(in, out, n, n_dev, temp, total);
	}
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
