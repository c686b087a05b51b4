// Shared definitions for the gfx950 simplifier kernels and their host launchers.
//
// Every device file is compiled with -ffp-contract=off (and the pragma below as a second
// guard): the quadric sums must round exactly like the reference's x86 build, and a fused
// multiply-add would change the collapse order.
#pragma once

#include <hip/hip_runtime.h>
#include <float.h>
#include <stddef.h>
#include <stdint.h>

#pragma clang fp contract(off)

namespace mhip
{

constexpr uint32_t kNone = 0xffffffffu;
constexpr int kMaxAttributes = 32;
// head rounds of a pass's commit (launched on the device's own list length), and the slots that
// record their list sizes (the window's heads and two for the extension)
constexpr int kMaxHeadRounds = 256;
constexpr int kHeadSizeSlots = kMaxHeadRounds + 2;
// the commit rounds' work list is split into one segment per XCD (8 on the MI355X), by the range
// of each rank's source class; workgroup i of a round works on segment i mod 8
constexpr int kListSegments = 8;

// vertex kinds (src/simplifier.cpp:296-305)
enum : uint8_t
{
	kManifold = 0,
	kBorder = 1,
	kSeam = 2,
	kComplex = 3,
	kLocked = 4,
};

// option bits (src/meshoptimizer.h:466-482, src/simplifier.cpp:2337-2342)
enum : uint32_t
{
	kOptLockBorder = 1u << 0,
	kOptSparse = 1u << 1,
	kOptErrorAbsolute = 1u << 2,
	kOptPrune = 1u << 3,
	kOptRegularize = 1u << 4,
	kOptPermissive = 1u << 5,
	kOptRegularizeLight = 1u << 6,
	kOptInternalSolve = 1u << 29,
	kOptInternalDebug = 1u << 30,
};

// vertex_lock flags (src/meshoptimizer.h:487-495)
enum : uint8_t
{
	kVertexLock = 1,
	kVertexProtect = 2,
	kVertexPriority = 4,
};

// commit status of a ranked candidate during the priority rounds
enum : uint8_t
{
	kUndecided = 0,
	kSkipped = 1,
	kRejected = 2,
	kCommitted = 3,
};

// 11-float quadric padded to 48 bytes so a gather is three 16-byte loads
struct __align__(16) Quad
{
	float a00, a11, a22, a10;
	float a20, a21, b0, b1;
	float b2, c, w, pad;
};

struct __align__(16) Grad
{
	float gx, gy, gz, gw;
};

// attributes with weight > 0: caller's attribute index and weight of each
struct AttrMap
{
	uint32_t index[kMaxAttributes];
	float weight[kMaxAttributes];
};

// scalars of one simplify call that live on the device
struct CallScalars
{
	float lo[3];          // bounding box minimum (rescale offset)
	float scale;          // 1 / extent (0 for a flat mesh)
	float extent;         // largest bounding box side
	float result_error;   // squared error so far (prefix max of committed collapses)
	float vertex_error;   // squared positional error (attribute path)
	uint32_t vertex_error_bits; // atomicMax target for vertex_error (non-negative floats order as uints)
	uint64_t box_min[3];  // (ordered key << 32 | index) reductions for the bounding box
	uint64_t box_max[3];
	uint32_t candidates;  // C_p of the current pass
	uint32_t triangles;   // live triangles T_p
	uint32_t next_triangles;
	uint32_t capacity_tris; // triangles processed by pick before the capacity cut
	uint32_t committed;   // commits below the stop rank
	uint32_t stop;        // stop rank of the pass
	uint32_t stop_reason; // 0 out of collapses, 1 error limit, 2 count limit, 3 error goal
	uint32_t skipped;
	uint32_t rejected;
	uint32_t tri_collapses;
	uint32_t list_count[3][kListSegments]; // undecided work-list counters of the rounds (rotating), per segment
	uint32_t list_off[kListSegments];      // where each segment starts in the list buffers
	uint32_t window_count[kListSegments];  // ranks per segment of the window being opened (zero otherwise)
	uint32_t dual_count;  // boundEdgeCollapses reduction
	uint32_t unique_count; // sparse remap: number of referenced vertices
	uint32_t tail_rounds;   // rounds run by the single-workgroup tail of the commit
	uint32_t component_count; // Prune: connected components of the welded mesh
	uint32_t prune_min_bits;  // Prune: smallest component error among kept triangles (float bits)
	uint32_t prune_max_bits;  // Prune: largest component error at or below a cutoff (float bits)
	uint32_t head_sizes[kHeadSizeSlots]; // work-list size at the start of each head round
	uint32_t x_visits;      // ranks listed over every k_round_x launch of the call (roofline accounting)
	uint32_t x_launches;
	uint32_t round;         // the commit rounds' base of the pass (list buffers rotate with the round number)
	uint32_t epoch;         // the round tag before the base round (round minima are tagged, monotonic over the call)
	uint32_t extended;      // 1 when the pass's window was extended on the device to the whole list
	uint32_t pick_capacity; // boundEdgeCollapses (src/simplifier.cpp:1292-1308): pick stops before exceeding it
	uint32_t work[8];     // small counters of setup kernels
};

#define HIP_CHECK(expr)                                                                               \
	do                                                                                                \
	{                                                                                                 \
		hipError_t err_ = (expr);                                                                     \
		if (err_ != hipSuccess)                                                                       \
			::mhip::fatal_hip(err_, #expr, __FILE__, __LINE__);                                       \
	} while (0)

[[noreturn]] void fatal_hip(hipError_t err, const char* expr, const char* file, int line);

__device__ __forceinline__ uint32_t float_bits(float f)
{
	return __float_as_uint(f);
}

// order-preserving map of a float to uint32 with +0 and -0 mapped to the same key
__device__ __forceinline__ uint32_t ordered_key(float f)
{
	uint32_t u = __float_as_uint(f);
	if (u == 0x80000000u)
		u = 0;
	return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// 64-bit minimum / maximum over the 64 lanes of a wavefront
__device__ __forceinline__ uint64_t wave_min64(uint64_t v)
{
#pragma unroll
	for (int o = 32; o > 0; o >>= 1)
	{
		uint64_t t = __shfl_xor(v, o, 64);
		v = t < v ? t : v;
	}
	return v;
}

__device__ __forceinline__ uint64_t wave_max64(uint64_t v)
{
#pragma unroll
	for (int o = 32; o > 0; o >>= 1)
	{
		uint64_t t = __shfl_xor(v, o, 64);
		v = t > v ? t : v;
	}
	return v;
}

inline unsigned int div_up(size_t a, size_t b)
{
	return (unsigned int)((a + b - 1) / b);
}

} // namespace mhip
