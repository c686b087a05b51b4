// One clusterlod level on the device (demo/clusterlod.h:640-690): boundary locks over the level's
// groups, each group packed with its vertices in first-use order (what Sparse does inside the
// simplifier), the batched simplification of the packed groups, and the results mapped back.
#pragma once

#include "common.h"

namespace mhip
{

// groups of at most lds_slots / 2 indices are packed by one kernel in LDS (8 bytes per slot; a
// launch's lds_slots fits its largest group, up to kClodLdsSlotsMax, which holds every group the
// batched kernel's classes take); larger groups use hash slices in global memory
constexpr uint32_t kClodLdsSlotsMax = 12288;
// the LDS slot capacity of a level whose largest group has max_group_indices indices
uint32_t clod_lds_slots(uint32_t max_group_indices);

// device buffers of one level (carved by the host); I = indices of all groups, G = groups,
// V = vertices of the shared buffer
struct ClodBuffers
{
	const uint32_t* indices;       // I: every group's index buffer (shared vertex ids), concatenated
	const uint32_t* index_offsets; // G + 1
	const uint32_t* remap;         // V: position remap (lowest vertex with the same position)
	uint32_t group_count;
	uint32_t vertex_count;
	uint32_t lds_slots;            // LDS hash capacity of this launch (clod_lds_slots)

	uint32_t* first_group;         // V: lowest group that uses each remapped vertex (kNone: none)
	uint32_t* last_group;          // V: highest group that uses it (0: none)
	uint32_t* slot_key;            // 2I: hash slices of the large groups (at 2 * index_offsets[g]): vertex + 1, 0 = empty
	uint32_t* slot_first;          // 2I: first position of the vertex in its group
	uint32_t* slot_local;          // 2I: its packed id
	uint32_t* staged;              // I: group g's first-use vertices in order, at index_offsets[g]
	uint32_t* unique;              // G + 1: vertices per group, then their exclusive scan (packed vertex offsets)
	uint32_t* order;               // I: packed vertex -> shared vertex id
	uint32_t* packed_indices;      // I
	float* packed_positions;       // 3I
	uint8_t* packed_locks;         // I
	uint32_t* packed_result;       // I
	float* packed_attributes;      // I * attribute_count: each packed vertex's attributes (null: none)
	uint32_t attribute_count;
};

// per group: first-use ranking (packed_indices, staged, unique counts) and the first/last group of
// every used remapped vertex; needs first_group = kNone, last_group = 0 and, when large_groups,
// slot_key = 0 and slot_first = kNone beforehand
void launch_clod_group(hipStream_t s, const ClodBuffers& b, bool large_groups);
// locks[i] = (two or more groups use remap[i] ? Lock : 0) | base_locks[i] (base_locks may be null)
// (lockBoundary, demo/clusterlod.h:372-419, with the caller's protect and lock bits as base_locks)
void launch_clod_locks(hipStream_t s, const ClodBuffers& b, const uint8_t* base_locks, uint8_t* locks);
// after the host scans `unique`: the packed vertex list, positions and locks of every group
void launch_clod_place(hipStream_t s, const ClodBuffers& b, const float* positions, size_t stride_floats, const uint8_t* locks,
                       const float* attributes, size_t attribute_stride_floats);
// destination[index_offsets[g] + k] = order[unique[g] + packed_result[index_offsets[g] + k]], k < counts[g]
void launch_clod_unpack(hipStream_t s, const ClodBuffers& b, const uint32_t* counts, uint32_t* destination);

// the fallbacks of clodBuild's per-group simplify (demo/clusterlod.h:461-518): gather the packed
// data of the groups `list` (n of them) at a sub-batch's offsets, and put the sub-batch's results
// back in place of theirs
void launch_clod_gather(hipStream_t s, const ClodBuffers& b, const uint32_t* list, uint32_t n, const uint32_t* sub_io, const uint32_t* sub_vo,
                        uint32_t* sub_idx, float* sub_pos, uint8_t* sub_locks, float* sub_attr);
void launch_clod_scatter(hipStream_t s, const ClodBuffers& b, const uint32_t* list, uint32_t n, const uint32_t* sub_io, const uint32_t* sub_result,
                         const uint32_t* sub_counts, const float* sub_errors, uint32_t* counts, float* errors);
// simplifyFallback's deindexed subset of group g (n indices: one vertex per index) and the result
// of the sloppy call over it mapped back to the group's packed ids
void launch_clod_deindex(hipStream_t s, const ClodBuffers& b, uint32_t g, uint32_t n, float* subset_pos, uint8_t* subset_locks, uint32_t* subset_idx);
void launch_clod_sloppy_result(hipStream_t s, const ClodBuffers& b, uint32_t g, const uint32_t* lod, uint32_t count, uint32_t* counts, float* errors,
                               float error);
// every group's error limited by its input triangles' edge lengths (groups whose target exceeds
// their size excepted)
void launch_clod_edge_limit(hipStream_t s, const ClodBuffers& b, const uint32_t* targets, float limit, float* errors);
// the groups `list` (n of them) keep their own indices, error 0
void launch_clod_keep(hipStream_t s, const ClodBuffers& b, const uint32_t* list, uint32_t n, uint32_t* counts, float* errors);

} // namespace mhip
