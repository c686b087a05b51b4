// meshopt_simplifySloppy on the device (src/simplifier.cpp:2647-2778): grid quantisation with a
// guided search of the grid size, first-occurrence cell numbering, ordered cell quadrics, the
// lowest-error vertex per cell and first-occurrence triangle deduplication.
#pragma once

#include "common.h"

namespace mhip
{

// device buffers of one sloppy call (carved by the host from the context arena)
struct SloppyBuffers
{
	const uint32_t* indices; // I
	const float4* pos;       // V rescaled positions (x, y, z, 0)
	const uint8_t* lock;     // V vertex_lock or null
	uint32_t vertex_count;
	uint32_t index_count;

	uint32_t* count;         // 1 word: triangles counted by launch_sloppy_count
	uint32_t* ids;           // V cell ids of the chosen grid
	uint32_t* vtable;        // vtable_size: first vertex of each id (open addressing)
	uint32_t vtable_size;    // power of two
	uint32_t* rep;           // V: first vertex with the same id
	uint32_t* cell_scan;     // V + 1: rep flags, then their exclusive scan (cell numbers)
	uint32_t* cells;         // V: cell of each vertex
	uint32_t* cell_first;    // V: first vertex of each cell
	uint32_t* entry_keys;    // 3T (+ ping-pong halves below): cell of each (triangle, corner) entry
	uint32_t* entry_vals;    // 3T: entry ids 3t + k
	uint32_t* entry_keys_alt;
	uint32_t* entry_vals_alt;
	uint32_t* cell_off;      // V + 2: start of each cell's run in the sorted entries
	uint32_t* sort_temp;
	uint32_t* scan_temp;
	Quad* cq;                // V: cell quadrics
	Quad* face;              // T: face quadric of each triangle (weight 3 inside one cell)
	uint64_t* best;          // V: (error bits << 32 | vertex) minimum per cell
	uint32_t* cell_remap;    // V
	uint32_t* error_bits;    // 1 word: largest cell error (bits of a non-negative float)
	uint4* tri;              // T: canonical (a, b, c) of each triangle, w = 1 when its cells differ
	uint32_t* ttable;        // ttable_size: first triangle of each canonical triple
	uint32_t ttable_size;    // power of two
	uint32_t* keep;          // T + 1: first occurrences, then their exclusive scan
};

// triangles whose three cell ids differ at `grid_size` (countTriangles of computeVertexIds), into *count
void launch_sloppy_count(hipStream_t s, const SloppyBuffers& b, int grid_size);

// everything after the search at the chosen grid: writes the kept triangles to destination and
// leaves their index count in cell_scan / keep totals read by the host (see host.cpp)
void launch_sloppy_cells(hipStream_t s, const SloppyBuffers& b, int grid_size, uint32_t* cell_count_dev);
void launch_sloppy_quadrics(hipStream_t s, const SloppyBuffers& b, uint32_t cell_count);
void launch_sloppy_filter(hipStream_t s, const SloppyBuffers& b, uint32_t cell_count, uint32_t* destination, uint32_t* kept_dev);

} // namespace mhip
