// Host-side launchers for the simplifier's device kernels (setup.hip, pass.hip).
#pragma once

#include "common.h"

namespace mhip
{

// optional HIP-event timing of the round kernel k_round_x (the dominant kernel)
struct KernelTimer
{
	hipEvent_t* begin;   // event pairs recorded around each launch
	hipEvent_t* end;
	int capacity;
	int count;
};

// device state of one simplify call; every pointer is carved from one arena
struct Mesh
{
	KernelTimer* timer;
	uint32_t vertex_count;   // V (after the optional sparse remap)
	uint32_t index_count;    // I of the input
	uint32_t attr_count;     // attributes with weight > 0
	uint32_t capacity;       // collapse capacity (boundEdgeCollapses)

	float4* praw;            // input positions (x, y, z, 0)
	float4* pos;             // rescaled positions
	uint32_t* remap;         // position remap: lowest index with equal xyz
	uint32_t* wedge;         // cyclic wedge list
	uint8_t* kind;
	uint32_t* loop[2];       // loop / loopback (edge loops), ping-pong for remapEdgeLoops
	uint32_t* loopback[2];
	int loop_cur;

	Quad* vq;                // vertex quadrics (primaries)
	Grad* vg;                // volume gradients (primaries; the position solve with attributes only)
	Quad* aq;                // attribute quadrics (per wedge)
	Grad* ag;                // attribute gradients (V x A)
	float* va;               // weighted attributes (V x A)

	uint32_t* cremap;        // collapse_remap
	uint32_t* src_off;       // V + 1: candidate ranks grouped by source class
	uint32_t* src_rank;      // candidates
	uint32_t* src_val;       // the value each listed candidate's commit writes at the class primary
	uint32_t* lock_rank;     // rank of the commit that locked a class (collapse_locked)
	uint64_t* pend[2];       // per-class round minima of the commit rounds, two per class (16 B):
	                         // over ranks whose source is the class, and whose target is
	uint32_t epoch;

	uint32_t* adj_off;       // V + 1: welded corner CSR offsets of the current pass
	uint2* adj;              // I: (next, prev) per corner

	uint32_t* ib[2];         // index buffer ping-pong
	int ib_cur;

	uint32_t* tri_cnt;       // per-triangle counts (pick / remap compaction)
	uint32_t* tri_off;

	uint32_t* c_v0;          // candidates in pick order
	uint32_t* c_v1;
	float* c_err;
	uint32_t* keys;          // sort buffers
	uint32_t* vals;
	uint32_t* keys_alt;
	uint32_t* vals_alt;

	uint32_t* r_v0;          // candidates in rank order
	uint32_t* r_v1;
	uint32_t* r_r0;          // remap[r_v0]: source class
	uint32_t* r_r1;          // remap[r_v1]: target class
	float* r_err;
	uint8_t* status;
	uint4* list[2];          // undecided work lists: (rank, source class, target class, 0)
	uint32_t* stop_tile;     // per-tile aggregates of the stop scan (3 words each)

	uint32_t* temp;          // scan / sort temporaries
	CallScalars* sc;
};

// ---------------------------------------------------------------- setup (setup.hip)
void launch_check_indices(hipStream_t s, const uint32_t* ib, size_t index_count, size_t vertex_count, uint32_t* flag);
void launch_sparse_remap(hipStream_t s, const uint32_t* indices_in, uint32_t* indices_out, size_t index_count, size_t vertex_count,
                         uint32_t* first, uint32_t* flags, uint32_t* dense, uint32_t* sparse_remap, uint32_t* unique_count, uint32_t* temp);
void launch_gather_positions(hipStream_t s, const float* positions, size_t stride_floats, const uint32_t* sparse_remap, float4* out, size_t vertex_count);
void launch_position_remap(hipStream_t s, Mesh& m, uint32_t* table, size_t table_size);
void launch_wedges(hipStream_t s, Mesh& m, uint32_t* flags, uint32_t* scratch_a, uint32_t* scratch_b, uint32_t* scratch_c, uint32_t* scratch_d);
void launch_corner_csr(hipStream_t s, const uint32_t* ib, size_t index_count, const uint32_t* remap, size_t vertex_count,
                       uint32_t* offsets, uint32_t* keys, uint32_t* corners, uint32_t* keys_alt, uint32_t* vals_alt, uint32_t* temp);
void launch_classify(hipStream_t s, Mesh& m, const uint32_t* raw_off, const uint32_t* raw_corners, const uint8_t* vertex_lock,
                     const uint32_t* sparse_remap, uint32_t options);
void launch_rescale(hipStream_t s, Mesh& m);
void launch_extent_only(hipStream_t s, Mesh& m);
void launch_rescale_attributes(hipStream_t s, Mesh& m, const float* attributes, size_t stride_floats, const uint32_t* sparse_remap,
                               const uint32_t* attr_index, const float* attr_weight);
void launch_vertex_quadrics(hipStream_t s, Mesh& m, const uint32_t* welded_off, const uint32_t* welded_corners, const uint8_t* vertex_lock,
                            const uint32_t* sparse_remap, uint32_t options);
void launch_attribute_quadrics(hipStream_t s, Mesh& m, const uint32_t* raw_off, const uint32_t* raw_corners);
void launch_dual_count(hipStream_t s, Mesh& m, const uint32_t* raw_off);

// ---------------------------------------------------------------- passes (pass.hip)
// tri_count sizes the grids; `live`, when given, is a device word holding the actual count (at
// most tri_count), so a pass can be enqueued before the host has read it
void launch_adjacency(hipStream_t s, Mesh& m, size_t tri_count, const uint32_t* live = nullptr);
void launch_pick(hipStream_t s, Mesh& m, size_t tri_count, const uint32_t* live = nullptr);
void launch_rank_sort(hipStream_t s, Mesh& m, size_t cand_count);
// commit state of a pass; with source_lists, also the candidate ranks grouped by source class
// that deep rounds read
void launch_perform_init(hipStream_t s, Mesh& m, size_t cand_count, bool source_lists);
void launch_source_lists(hipStream_t s, Mesh& m, size_t cand_count);
void launch_pick_capacity(hipStream_t s, Mesh& m, size_t index_count);
void launch_round(hipStream_t s, Mesh& m, size_t list_bound, int head_index, int rel, bool deep);
void launch_round_tail(hipStream_t s, Mesh& m, int rel, uint32_t max_rounds, bool deep);
void launch_window_extend(hipStream_t s, Mesh& m, size_t lo, size_t hi, bool if_no_stop);
void launch_resolve(hipStream_t s, Mesh& m, size_t lo, size_t hi);
void launch_stop_scan(hipStream_t s, Mesh& m, size_t window, size_t cand_count, uint32_t goal, float error_limit, bool conditional);
void launch_stop_commit(hipStream_t s, Mesh& m, size_t window, uint32_t goal, float error_limit);
void launch_apply(hipStream_t s, Mesh& m, size_t cand_count);
void launch_remap_triangles(hipStream_t s, Mesh& m, size_t tri_count);
void launch_debug_bits(hipStream_t s, Mesh& m, size_t index_count);
void launch_sparse_backmap(hipStream_t s, uint32_t* ib, size_t index_count, const uint32_t* sparse_remap);

// ---------------------------------------------------------------- vertex solve (solve.hip)
// buffers of meshopt_simplifyWithUpdate's solve, carved next to the mesh
struct SolveBuffers
{
	uint8_t* update;   // V: vertex of the result, or primary of one
	uint32_t* listed;  // V: 1 for the primaries the position solve visits
	uint32_t* waiting; // V: listed lower-index neighbour entries not solved yet
	uint32_t* queue[2]; // V each: primaries ready in the current / next round
	float4* cand;      // V: a listed primary's solved position (w unused)
	uint8_t* state;    // V: its outcome as far as decided (open, keep, move)
};
// the caller's arrays written back (finalizeVertices); strides in floats
struct SolveOutput
{
	float* positions;
	size_t position_stride;
	float* attributes;
	size_t attribute_stride;
	const uint8_t* vertex_lock;
	const uint32_t* sparse_remap;
	AttrMap map;
};
// marks the result's vertices, rebuilds the welded adjacency of the result, lists the
// primaries to solve (sc->work[4] = how many) and queues those with no listed lower-index
// neighbour for the first round; sc->work[5 + r % 3] holds round r's queue length
void launch_solve_setup(hipStream_t s, Mesh& m, const SolveBuffers& b, size_t result_count);
// `rounds` rounds of the ordered position solve starting at round `first` (0-based); the solve
// is complete once the queue of the next round to run is empty; `listed` sizes the grid
void launch_solve_rounds(hipStream_t s, Mesh& m, const SolveBuffers& b, uint32_t listed, uint32_t first, uint32_t rounds);
// wedge copies, attribute solve and the write-back into the caller's arrays
void launch_solve_finish(hipStream_t s, Mesh& m, const SolveBuffers& b, const SolveOutput& o);

// ---------------------------------------------------------------- Prune (prune.hip)
// components of the welded mesh over the input index buffer: comp[V] (ids in order of each
// component's lowest primary), sc->component_count; scratch: five arrays of V + 1 words
void launch_components(hipStream_t s, Mesh& m, size_t index_count, uint32_t* comp, uint32_t* scratch[5]);
// squared bounding-sphere radius per component around its vertex mean, sums in vertex order;
// sc->prune_min_bits = the smallest (FLT_MAX bits when there is none)
void launch_measure_components(hipStream_t s, Mesh& m, const uint32_t* comp, float* cerr, uint32_t* scratch[5]);
// keep the triangles whose component error (of their first vertex) is above cutoff, in order;
// sc->next_triangles and sc->prune_min_bits (over the kept ones) are the results
void launch_prune_triangles(hipStream_t s, Mesh& m, size_t tri_count, const uint32_t* comp, const float* cerr, float cutoff);
// sc->prune_max_bits = largest component error <= cutoff (0 if none)
void launch_component_max(hipStream_t s, Mesh& m, const float* cerr, float cutoff);

} // namespace mhip
