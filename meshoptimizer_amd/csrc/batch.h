// Batched simplification of many small meshes, one workgroup per mesh with LDS-resident state.
#pragma once

#include "common.h"

namespace mhip
{

// LDS size classes of one batched call (launched side by side)
constexpr int kBatchClasses = 4;
// the class that a mesh overflowing the candidate arrays of classes 0-2 runs again in (the last
// class, one workgroup per CU, holds the full bound)
constexpr int kBatchRetryClass = 3;

struct BatchArgs
{
	uint32_t* destination;             // packed: mesh p writes at index_offsets[p]
	uint32_t* result_counts;           // per mesh
	float* result_errors;              // per mesh (may be null)
	uint32_t* status;                  // per mesh: 0 done, 1 plain path, 2 bad index, 3 pass bound hit (internal error), 4 invalid sizes,
	                                   // 5 more candidates than the layout holds (run again in kBatchRetryClass),
	                                   // 6 waiting for the last class's launch
	uint32_t* stats;                   // 4 words per mesh (may be null): passes, and summed over them the live triangles at
	                                   // the pass start, the candidates and the committed collapses
	const uint32_t* indices;           // packed local indices
	const uint32_t* index_offsets;     // mesh_count + 1
	const float* positions;            // packed vertices (position in the first 12 bytes)
	const uint32_t* vertex_offsets;    // mesh_count + 1
	const uint8_t* vertex_lock;        // per packed vertex (like the positions), or null
	size_t vertex_count;               // vertices the positions hold; every vertex_offsets[p + 1] is at most this
	Quad* quadrics;                    // vertex_count quadrics: mesh p keeps its vertex quadrics at vertex_offsets[p]
	size_t position_stride_floats;
	const uint32_t* target_index_counts; // per mesh
	size_t patch_count;
	float target_error;
	uint32_t options;
	// attributes (attribute_count 0: none): packed per vertex like the positions; the attributes
	// with weight > 0 (attr_map) enter the quadrics, rescaled into attr_values; the per-vertex
	// attribute quadrics and gradients live in global memory beside the vertex quadrics (mesh p
	// at vertex_offsets[p], like them)
	const float* attributes;
	size_t attribute_stride_floats;
	uint32_t attribute_count;
	AttrMap attr_map;
	float* attr_values;   // vertex_count * attribute_count
	Quad* attr_quadrics;  // vertex_count
	Grad* attr_grads;     // vertex_count * attribute_count
};

// The launch shape of the size classes: per class, the workgroups that share one CU (the class's
// LDS share is the CU's 160 KB over that count), the threads per workgroup (128, 256, 512 or 1024), and
// whether its workgroups take meshes from a cursor (grid: per_cu workgroups per CU) instead of one
// workgroup per mesh.  The last class always runs from a cursor and launches after the status
// read-back (unless the host knows a mesh in its range).  The default comes from the LDS budget;
// a measured one (meshoptimizer_amd/tune.py) may replace it.
struct BatchShape
{
	uint32_t per_cu[kBatchClasses];
	uint32_t threads[kBatchClasses];
	uint32_t cursor[kBatchClasses];
};

BatchShape batch_default_shape();
// null when the shape can run: widths of 128/256/512/1024, at most 6 wavefronts per SIMD for the
// 80-VGPR widths (4 for 1024), workgroups per CU non-increasing over the classes, the last class
// one workgroup per CU; otherwise the reason
const char* batch_shape_error(const BatchShape& shape);

// size classes: a mesh belongs to the first class whose vertex and index caps it fits; ccap is the
// class's candidate capacity (class 3 holds the full bound, so it never returns status 5)
struct BatchClasses
{
	uint32_t vcap[kBatchClasses];
	uint32_t icap[kBatchClasses];
	uint32_t ccap[kBatchClasses];
	BatchShape shape;
};

// the largest caps per class whose LDS layout fits the class's share of a CU, given the LDS one
// workgroup may take (a class that the limit leaves no larger than the one before it is empty:
// caps 0)
BatchClasses batch_classes(uint32_t max_lds_per_workgroup, const BatchShape& shape);
// LDS bytes of a layout at a width; lets both launch forms of a width take `bytes` of dynamic LDS
// (false when the runtime refuses)
size_t batch_lds_bytes(uint32_t threads, uint32_t vmax, uint32_t imax, uint32_t cmax);
bool batch_allow_lds(uint32_t threads, size_t bytes);

// per mesh: validate sizes (status 4), pick its class (ids[c * patch_count + slot], counts[c])
// or mark it for the plain path (status 1); counts must be zero on entry
void launch_batch_classify(hipStream_t s, const BatchArgs& a, const BatchClasses& classes, uint32_t* ids, uint32_t* counts);
// host_counts (may be null): a host-visible copy of the class counts
void launch_batch_order(hipStream_t s, const BatchArgs& a, uint32_t* ids, const uint32_t* counts, uint32_t* temp, uint32_t* host_counts);
// batches of up to this many meshes are classified and ordered in one launch
constexpr size_t kFusedOrderMax = 16384;
// classify and order in one launch (counts written, no clearing needed); false when the batch is
// larger than kFusedOrderMax (then launch_batch_classify and launch_batch_order)
bool launch_batch_classify_order(hipStream_t s, const BatchArgs& a, const BatchClasses& classes, uint32_t* ids, uint32_t* counts, uint32_t* host_counts);

// the meshes of one class (ids, *count_dev of them), every one with the (vmax, imax, cmax) layout,
// on workgroups of `threads`: cursor == null gives workgroup b mesh b (grid at least the count;
// the workgroups past it leave at once); otherwise `grid` workgroups take meshes in list order
// through the cursor word (cleared here on the stream) until the list ends
void launch_batch(hipStream_t s, uint32_t threads, const BatchArgs& a, uint32_t vmax, uint32_t imax, uint32_t cmax, const uint32_t* ids,
                  const uint32_t* count_dev, size_t grid, uint32_t* cursor);

} // namespace mhip
