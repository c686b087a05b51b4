// Batched simplification of many small meshes, one workgroup per mesh with LDS-resident state.
#pragma once

#include "common.h"

namespace mhip
{

struct BatchArgs
{
	uint32_t* destination;             // packed: mesh p writes at index_offsets[p]
	uint32_t* result_counts;           // per mesh
	float* result_errors;              // per mesh (may be null)
	uint32_t* status;                  // per mesh: 0 done, 1 too large for LDS, 2 bad index, 3 pass bound hit (internal error)
	uint32_t* passes;                  // per mesh (may be null)
	const uint32_t* indices;           // packed local indices
	const uint32_t* index_offsets;     // mesh_count + 1
	const float* positions;            // packed vertices (position in the first 12 bytes)
	const uint32_t* vertex_offsets;    // mesh_count + 1
	size_t position_stride_floats;
	const uint32_t* target_index_counts; // per mesh
	size_t patch_count;
	float target_error;
	uint32_t options;
};

// LDS size classes of one batched call (launched side by side)
constexpr int kBatchClasses = 3;

size_t batch_lds_bytes(uint32_t vmax, uint32_t imax);
// one workgroup per listed mesh (ids: count device words), every one with the (vmax, imax) layout
void launch_batch(hipStream_t s, const BatchArgs& a, uint32_t vmax, uint32_t imax, const uint32_t* ids, size_t count);

} // namespace mhip
