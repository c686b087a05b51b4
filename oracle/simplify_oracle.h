/*
 * simplify_oracle.h -- CPU restatement of meshoptimizer 1.2's edge-collapse simplifier.
 * TEST INFRASTRUCTURE ONLY (the checker for the HIP path; never linked into the product).
 */
#ifndef SIMPLIFY_ORACLE_H
#define SIMPLIFY_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MORACLE_MAX_PASSES 64

typedef struct
{
	int perform_rounds;  /* 0: sequential greedy (src/simplifier.cpp:1498); 1: round-based commit */
	size_t round_window; /* initial rank window for the round-based commit; 0 = all candidates */
	int solve_rounds;    /* meshopt_simplifyWithUpdate's position solve: 0 in index order, 1 in dependency rounds */
} moracle_config;

typedef struct
{
	unsigned int candidates, evaluated, committed, skipped, rejected;
	unsigned int rounds, window, stop, tri_collapses, live_triangles;
} moracle_pass_stats;

typedef struct
{
	int passes;
	unsigned int unique_positions;
	unsigned int kinds[5];
	float result_error_sq;
	float vertex_error_sq;
	unsigned int solve_rounds; /* rounds of the position solve when run in rounds */
	moracle_pass_stats pass[MORACLE_MAX_PASSES];
} moracle_stats;

size_t moracle_simplify_edge(unsigned int* destination, const unsigned int* indices, size_t index_count, const float* vertex_positions, size_t vertex_count, size_t vertex_positions_stride, const float* vertex_attributes, size_t vertex_attributes_stride, const float* attribute_weights, size_t attribute_count, const unsigned char* vertex_lock, size_t target_index_count, float target_error, unsigned int options, float* result_error, const moracle_config* config, moracle_stats* stats);

size_t moracle_simplify(unsigned int* destination, const unsigned int* indices, size_t index_count, const float* vertex_positions, size_t vertex_count, size_t vertex_positions_stride, size_t target_index_count, float target_error, unsigned int options, float* result_error);

size_t moracle_simplify_with_attributes(unsigned int* destination, const unsigned int* indices, size_t index_count, const float* vertex_positions, size_t vertex_count, size_t vertex_positions_stride, const float* vertex_attributes, size_t vertex_attributes_stride, const float* attribute_weights, size_t attribute_count, const unsigned char* vertex_lock, size_t target_index_count, float target_error, unsigned int options, float* result_error);

/* meshopt_simplifyWithUpdate: indices, positions and attributes updated in place */
size_t moracle_simplify_with_update(unsigned int* indices, size_t index_count, float* vertex_positions, size_t vertex_count, size_t vertex_positions_stride, float* vertex_attributes, size_t vertex_attributes_stride, const float* attribute_weights, size_t attribute_count, const unsigned char* vertex_lock, size_t target_index_count, float target_error, unsigned int options, float* result_error);

float moracle_simplify_scale(const float* vertex_positions, size_t vertex_count, size_t vertex_positions_stride);

/* the grid-size search of meshopt_simplifySloppy, pass by pass (grid size tried, triangles counted) */
typedef struct
{
	int passes;
	int final_grid;
	unsigned int cells;
	int grid[MORACLE_MAX_PASSES];
	unsigned int triangles[MORACLE_MAX_PASSES];
} moracle_sloppy_trace;

size_t moracle_simplify_sloppy(unsigned int* destination, const unsigned int* indices, size_t index_count, const float* vertex_positions, size_t vertex_count, size_t vertex_positions_stride, const unsigned char* vertex_lock, size_t target_index_count, float target_error, float* result_error, moracle_sloppy_trace* trace);

#ifdef __cplusplus
}
#endif

#endif
