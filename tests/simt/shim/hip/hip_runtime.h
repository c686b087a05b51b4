// CPU stand-in for the parts of the HIP device language that the batched kernel's device code
// uses (tests/simt): each GPU thread of a workgroup runs as a fiber, __syncthreads is a
// workgroup rendezvous and every wavefront operation (shuffles, ballots, readlane, wave barrier)
// is a rendezvous of the 64 fibers of that wavefront.  Test infrastructure only: it lets the
// kernel's own source run on the host, where hangs, divergent barriers and out-of-range accesses
// can be found without a GPU.
#pragma once

#include <math.h>
#include <stdint.h>
#include <string.h>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __align__(n) alignas(n)

struct float4
{
	float x, y, z, w;
};
struct float2
{
	float x, y;
};
struct uint2
{
	unsigned int x, y;
};
struct uint4
{
	unsigned int x, y, z, w;
};

inline float4 make_float4(float x, float y, float z, float w) { return float4{x, y, z, w}; }

typedef int hipError_t;
typedef void* hipStream_t;
enum
{
	hipSuccess = 0
};

namespace simt
{
struct Dim
{
	unsigned int x, y, z;
};
const Dim& thread_idx();
void block_barrier(int line);
const uint64_t* wave_sync(uint64_t v, int line);
unsigned int lane();

template <class T>
inline uint64_t to_bits(T v)
{
	uint64_t u = 0;
	memcpy(&u, &v, sizeof(T));
	return u;
}

template <class T>
inline T from_bits(uint64_t u)
{
	T v;
	memcpy(&v, &u, sizeof(T));
	return v;
}

template <class T>
inline T shfl_up(T v, unsigned int delta, int line)
{
	const uint64_t* a = wave_sync(to_bits(v), line);
	unsigned int l = lane();
	return l >= delta ? from_bits<T>(a[l - delta]) : v;
}

template <class T>
inline T shfl_xor(T v, int mask, int line)
{
	const uint64_t* a = wave_sync(to_bits(v), line);
	return from_bits<T>(a[(lane() ^ (unsigned int)mask) & 63]);
}

inline uint64_t ballot(bool p, int line)
{
	const uint64_t* a = wave_sync(p ? 1 : 0, line);
	uint64_t m = 0;
	for (int l = 0; l < 64; ++l)
		m |= (a[l] ? 1ull : 0ull) << l;
	return m;
}

template <class T>
inline T readlane(T v, unsigned int j, int line)
{
	const uint64_t* a = wave_sync(to_bits(v), line);
	return from_bits<T>(a[j & 63]);
}
} // namespace simt

#define threadIdx (::simt::thread_idx())
#define __syncthreads() ::simt::block_barrier(__LINE__)
#define __shfl_up(v, d, w) ::simt::shfl_up((v), (d), __LINE__)
#define __shfl_xor(v, m, w) ::simt::shfl_xor((v), (m), __LINE__)
#define __ballot(p) ::simt::ballot((p), __LINE__)
#define __builtin_amdgcn_readlane(v, j) ::simt::readlane((v), (j), __LINE__)
// every fiber of the wavefront must reach it (all lanes are active in this model: lane 0 is first)
#define __builtin_amdgcn_readfirstlane(v) ::simt::readlane((v), 0u, __LINE__)
#define __builtin_amdgcn_wave_barrier() ((void)::simt::wave_sync(0, __LINE__))
#define __builtin_amdgcn_fence(...) ((void)0)
#define __lane_id() (::simt::lane())
#define __popcll(x) __builtin_popcountll(x)
#define __clzll(x) __builtin_clzll((unsigned long long)(x))
#define clock64() 0ull

inline uint32_t __float_as_uint(float f) { return ::simt::from_bits<uint32_t>(::simt::to_bits(f)); }
inline float __uint_as_float(uint32_t u) { return ::simt::from_bits<float>(::simt::to_bits(u)); }
inline int __float_as_int(float f) { return ::simt::from_bits<int>(::simt::to_bits(f)); }
inline float __int_as_float(int u) { return ::simt::from_bits<float>(::simt::to_bits(u)); }

// fibers switch only at rendezvous points, so a read-modify-write is atomic by construction
inline unsigned int atomicAdd(unsigned int* p, unsigned int v) { unsigned int o = *p; *p = o + v; return o; }
inline unsigned int atomicOr(unsigned int* p, unsigned int v) { unsigned int o = *p; *p = o | v; return o; }
inline unsigned int atomicMax(unsigned int* p, unsigned int v) { unsigned int o = *p; *p = o > v ? o : v; return o; }
inline unsigned int atomicMin(unsigned int* p, unsigned int v) { unsigned int o = *p; *p = o < v ? o : v; return o; }
inline unsigned int atomicCAS(unsigned int* p, unsigned int cmp, unsigned int v) { unsigned int o = *p; if (o == cmp) *p = v; return o; }
inline unsigned long long atomicMax(unsigned long long* p, unsigned long long v) { unsigned long long o = *p; *p = o > v ? o : v; return o; }
inline unsigned long long atomicMin(unsigned long long* p, unsigned long long v) { unsigned long long o = *p; *p = o < v ? o : v; return o; }
