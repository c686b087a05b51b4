/**
 * meshopt_hip.h -- entry points of libmeshopt_hip.so that have no counterpart in the
 * reference library: device-resident calls, per-pass statistics and error reporting.
 * Kept apart from meshoptimizer.h so that the drop-in header stays byte-compatible in shape.
 */
#ifndef MESHOPT_HIP_H
#define MESHOPT_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C"
{
#endif

/**
 * meshopt_simplifyWithAttributes on buffers that already live in device memory (HBM):
 * destination, indices, vertex_positions, vertex_attributes and vertex_lock are device
 * pointers; attribute_weights and result_error are host pointers.  The call is synchronous.
 * Semantics and result are those of meshopt_simplifyWithAttributes (src/meshoptimizer.h:537);
 * with vertex_attributes == NULL, attribute_count == 0 and vertex_lock == NULL it is
 * meshopt_simplify (src/meshoptimizer.h:513).
 * Returns the index count, or (size_t)-1 on a HIP error (see meshopt_hip_last_error).
 */
size_t meshopt_hip_simplify_device(unsigned int* destination, const unsigned int* indices, size_t index_count, const float* vertex_positions, size_t vertex_count, size_t vertex_positions_stride, const float* vertex_attributes, size_t vertex_attributes_stride, const float* attribute_weights, size_t attribute_count, const unsigned char* vertex_lock, size_t target_index_count, float target_error, unsigned int options, float* result_error);

/**
 * meshopt_simplifyWithUpdate on device-resident arrays: indices, vertex_positions,
 * vertex_attributes and vertex_lock are device pointers, updated in place as the host entry
 * updates its arrays; attribute_weights and result_error are host pointers.  Replaces the
 * reference's meshopt_simplifyWithUpdate (src/meshoptimizer.h:560) for callers whose buffers
 * live in HBM.  Returns the index count, or (size_t)-1 with meshopt_hip_last_error set.
 */
size_t meshopt_hip_simplify_with_update_device(unsigned int* indices, size_t index_count, float* vertex_positions, size_t vertex_count, size_t vertex_positions_stride, float* vertex_attributes, size_t vertex_attributes_stride, const float* attribute_weights, size_t attribute_count, const unsigned char* vertex_lock, size_t target_index_count, float target_error, unsigned int options, float* result_error);

/**
 * meshopt_simplifySloppy on device-resident arrays (every pointer but result_error is a device
 * pointer); no counterpart in the reference (its entry point takes host arrays).  Returns the
 * number of indices written, or (size_t)-1 with meshopt_hip_last_error set.
 */
size_t meshopt_hip_simplify_sloppy_device(unsigned int* destination, const unsigned int* indices, size_t index_count, const float* vertex_positions, size_t vertex_count, size_t vertex_positions_stride, const unsigned char* vertex_lock, size_t target_index_count, float target_error, float* result_error);

/**
 * The per-group fallbacks of clusterlod's simplify (demo/clusterlod.h:461-518), configured as
 * clodConfig's simplify_* fields (demo/clusterlod.h:42-53) configure them.
 */
typedef struct meshopt_hip_clod_config
{
	int fallback_permissive;   /* simplify_fallback_permissive: a group left above its target runs again with meshopt_SimplifyPermissive (only when options lack it) */
	int fallback_sloppy;       /* simplify_fallback_sloppy: a group still above its target takes meshopt_simplifySloppy over its deindexed triangles */
	float error_factor_sloppy; /* simplify_error_factor_sloppy: the sloppy result's absolute error is multiplied by this */
	float error_edge_limit;    /* simplify_error_edge_limit: when > 0, a group's error is limited by its triangles' edge lengths */
} meshopt_hip_clod_config;

/**
 * One clusterlod level on the device (the group loop of clodBuild, demo/clusterlod.h:640-690):
 * group g owns indices [index_offsets[g], index_offsets[g+1]) of `indices`, which reference the
 * shared vertex buffer (vertex_count vertices, remap = its position remap, as
 * meshopt_generatePositionRemap gives it).  locks[i] (vertex_count bytes, written) receives the
 * level's locks: meshopt_SimplifyVertex_Lock on every vertex whose position two or more groups use
 * (lockBoundary, demo/clusterlod.h:372-419), ORed with base_locks[i] (may be NULL: the caller's
 * protect and lock bits).  Every group is then simplified as clodBuild's simplify() does it
 * (demo/clusterlod.h:461-518): meshopt_simplifyWithAttributes(group indices, the shared vertices,
 * the shared attributes (vertex_attributes: device array of vertex_count rows of
 * vertex_attributes_stride bytes; attribute_weights: host array of attribute_count weights;
 * attribute_count 0: none), locks, target_index_counts[g], target_error,
 * options | meshopt_SimplifySparse) (clodBuild passes Sparse | ErrorAbsolute | Permissive and
 * FLT_MAX, demo/clusterlod.h:466-472), all groups in one batched
 * launch; with config (may be NULL: no fallbacks), the groups left above their target run again
 * with Permissive and then through meshopt_simplifySloppy on their deindexed triangles, and every
 * error is limited by edge length, as the config's fields say.  A group whose target exceeds its
 * index count keeps its indices with error 0.  Results: at destination + index_offsets[g] in
 * shared vertex ids, result_counts[g], result_errors[g] (may be NULL).  All arrays are device
 * pointers.  Returns 0, or -1 with meshopt_hip_last_error set.  No counterpart in the reference
 * library (clusterlod is its demo).
 */
int meshopt_hip_clod_level_device(unsigned int* destination, unsigned int* result_counts, float* result_errors, const unsigned int* indices, const unsigned int* index_offsets, size_t group_count, const float* vertex_positions, size_t vertex_count, size_t vertex_positions_stride, const float* vertex_attributes, size_t vertex_attributes_stride, const float* attribute_weights, size_t attribute_count, const unsigned int* remap, const unsigned char* base_locks, unsigned char* locks, const unsigned int* target_index_counts, float target_error, unsigned int options, const meshopt_hip_clod_config* config);

/**
 * Simplifies many independent meshes (for example the cluster groups of a clusterlod build) in
 * one launch; every mesh gets exactly what meshopt_simplifyWithAttributes (src/meshoptimizer.h:537)
 * returns for it with no attributes (the _with_attributes variants: with its slice of the packed
 * attributes, vertex_attributes_stride bytes per vertex, and the shared host array of
 * attribute_count weights), its slice of vertex_lock and the shared options.  Mesh p
 * owns indices [index_offsets[p], index_offsets[p+1]) (local to its vertices) and vertices
 * [vertex_offsets[p], vertex_offsets[p+1]) of the packed arrays, and writes its result at
 * destination + index_offsets[p]; result_counts[p] and result_errors[p] (may be NULL) receive
 * the index count and error.  vertex_lock (may be NULL) holds one meshopt_SimplifyVertex_* byte
 * per packed vertex.  target_error and options are shared by all meshes (Sparse and Prune
 * batches run mesh by mesh on the plain path; a clusterlod group packed with its vertices in
 * first-use order is what Sparse would make of it).  Returns 0, or -1 with meshopt_hip_last_error
 * set.  The _device variant takes device pointers for every array and the number of vertices the
 * positions hold (vertex_offsets[mesh_count] must not exceed it); the other copies host arrays in
 * and out around the same launch.  No counterpart in the reference, whose callers loop over
 * meshopt_simplify* per group (demo/clusterlod.h:470-479).
 */
int meshopt_hip_simplify_batch_device(unsigned int* destination, unsigned int* result_counts, float* result_errors, const unsigned int* indices, const unsigned int* index_offsets, const float* vertex_positions, const unsigned int* vertex_offsets, const unsigned char* vertex_lock, size_t vertex_count, size_t vertex_positions_stride, const unsigned int* target_index_counts, size_t mesh_count, float target_error, unsigned int options);
int meshopt_hip_simplify_batch(unsigned int* destination, unsigned int* result_counts, float* result_errors, const unsigned int* indices, const unsigned int* index_offsets, const float* vertex_positions, const unsigned int* vertex_offsets, const unsigned char* vertex_lock, size_t vertex_positions_stride, const unsigned int* target_index_counts, size_t mesh_count, float target_error, unsigned int options);
int meshopt_hip_simplify_batch_with_attributes_device(unsigned int* destination, unsigned int* result_counts, float* result_errors, const unsigned int* indices, const unsigned int* index_offsets, const float* vertex_positions, const unsigned int* vertex_offsets, const float* vertex_attributes, size_t vertex_attributes_stride, const float* attribute_weights, size_t attribute_count, const unsigned char* vertex_lock, size_t vertex_count, size_t vertex_positions_stride, const unsigned int* target_index_counts, size_t mesh_count, float target_error, unsigned int options);
int meshopt_hip_simplify_batch_with_attributes(unsigned int* destination, unsigned int* result_counts, float* result_errors, const unsigned int* indices, const unsigned int* index_offsets, const float* vertex_positions, const unsigned int* vertex_offsets, const float* vertex_attributes, size_t vertex_attributes_stride, const float* attribute_weights, size_t attribute_count, const unsigned char* vertex_lock, size_t vertex_positions_stride, const unsigned int* target_index_counts, size_t mesh_count, float target_error, unsigned int options);

/**
 * meshopt_hip_simplify_batch_device that allocates nothing: no counterpart in the reference; it is
 * for callers that cannot allocate device or host memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same arguments and results as
 * meshopt_hip_simplify_batch_device.  Its scratch (status words, size-class lists and the vertex
 * quadrics of the batch) is one device array of 64 MiB per GPU that the library holds until the
 * process ends; its host words are a static array.  Limits: at most 16384 meshes, at most
 * (64 MiB - the lists) / 48 vertices (about 1.39 million), every mesh within the batched kernel's
 * LDS size classes (a mesh that needs the plain path makes the call fail), and no Sparse or Prune
 * option; a call beyond them returns -1 with meshopt_hip_last_error set, before any launch for the
 * size limits.  Every call on one GPU shares that GPU's scratch: calls that overlap on one GPU race
 * (the library serialises calls made through it; callers that bypass its lock must not overlap).
 * The first call on a device creates the library's streams there and looks up the scratch array's
 * device address.
 */
int meshopt_hip_simplify_batch_device_noalloc(unsigned int* destination, unsigned int* result_counts, float* result_errors, const unsigned int* indices, const unsigned int* index_offsets, const float* vertex_positions, const unsigned int* vertex_offsets, const unsigned char* vertex_lock, size_t vertex_count, size_t vertex_positions_stride, const unsigned int* target_index_counts, size_t mesh_count, float target_error, unsigned int options);

/* Per-pass counters of the last call made on the calling thread; the same quantities the
 * reference prints under TRACE (src/simplifier.cpp:1625-1633). */
typedef struct meshopt_hip_pass_stats
{
	unsigned int live_triangles; /* T_p at the start of the pass */
	unsigned int candidates;     /* C_p, picked edge collapses */
	unsigned int evaluated;      /* ranks looked at by the greedy loop, the stopping one included */
	unsigned int committed;
	unsigned int skipped;        /* endpoint already locked by a commit */
	unsigned int rejected;       /* would flip a triangle */
	unsigned int stop;           /* 0 out of collapses, 1 error limit, 2 count limit, 3 error goal */
	unsigned int rounds;         /* priority rounds the GPU commit used */
	unsigned int tail_rounds;    /* of them, the rounds run by the single-workgroup tail */
} meshopt_hip_pass_stats;

typedef struct meshopt_hip_call_stats
{
	unsigned int passes;
	unsigned int vertex_count;     /* after the optional sparse remap */
	unsigned int unique_positions; /* primaries of the position remap */
	unsigned int kinds[5];         /* primaries per kind: manifold, border, seam, complex, locked */
	float result_error_sq;
	float vertex_error_sq;
	unsigned int solve_vertices; /* meshopt_simplifyWithUpdate: primaries left to the ordered rounds (the
	                                outcomes not decided from their neighbours' possible positions) */
	unsigned int solve_rounds;   /* rounds of the ordered position solve */
	unsigned int read_backs;     /* host synchronisations with the device during the call */
	unsigned int device_extensions; /* passes whose rank window the device extended to the whole list */
} meshopt_hip_call_stats;

/* Copies up to max_passes pass records of the calling thread's last call; returns the pass count. */
int meshopt_hip_last_call_stats(meshopt_hip_call_stats* call, meshopt_hip_pass_stats* passes, int max_passes);

/* Policy of the plain path's commit rounds (no counterpart in the reference; results never depend
 * on it).  A rank's flip test is decided early only where it cannot change any more; the cheap
 * test waits while any corner of the rank's neighbourhood may still move, the deep test also
 * decides when the outcome is the same for every way the lower candidates could move those
 * corners (fewer rounds, dearer ones).  A window runs deep rounds when the head-round count the
 * library learned for windows of its size exceeds deep_heads (-1: every window; below -1: the
 * default, 64), from head round deep_from_head on (default 8) and in its tail.  Applies to later
 * calls in the process; returns 0, or -1 with meshopt_hip_last_error set. */
int meshopt_hip_set_round_policy(int deep_heads, int deep_from_head);

/* Forgets the head-round counts the library learned per window size (no counterpart in the
 * reference; results never depend on them): later calls start as the process's first did, with
 * a read-back of the list length after each batch of head rounds of a window size not yet seen.
 * For callers that time or test first calls; returns 0. */
int meshopt_hip_reset_round_counts(void);

/* HIP-event timing of the dominant kernel (k_round_x, the commit rounds) on the library's
 * stream.  Each call returns the totals gathered since the previous call (summed kernel time in
 * ms, launches, ranks listed over those launches), resets them, and turns timing on or off. */
void meshopt_hip_kernel_timing(int enable, double* total_ms, size_t* launches, size_t* ranks);

/* HIP-event timing of the batched kernel (meshopt_hip_simplify_batch*), gathered while
 * meshopt_hip_kernel_timing has timing on: summed kernel time in ms, launches and meshes since
 * the previous call; resets the totals. */
void meshopt_hip_batch_kernel_timing(double* total_ms, size_t* launches, size_t* meshes);

/* Per-mesh statistics of the last batched call (meshopt_hip_simplify_batch,
 * meshopt_hip_simplify_batch_device, meshopt_hip_clod_level_device): 4 words per mesh, the
 * passes, and summed over them the live triangles at each pass start, the candidates and the
 * committed collapses (the per-pass counters of src/simplifier.cpp:1629-1632).  Copies
 * min(max_meshes, meshes of that call) entries; returns their count, or -1 on a HIP error. */
int meshopt_hip_last_batch_stats(unsigned int* stats, size_t max_meshes);

/* Launch shape of the batched calls' four LDS size classes (no counterpart in the reference):
 * per class, the workgroups that share one CU (the class's LDS share is the CU's 160 KB over that
 * count, which sets the largest mesh it holds), the threads per workgroup (128, 256, 512 or 1024) and
 * whether its workgroups loop over the class's meshes through a cursor (cursor != 0) rather than
 * taking one mesh each.  The default follows from the LDS budget; at the first batched call the
 * library reads a measured shape from the table that meshoptimizer_amd/tune.py writes
 * (batch_shape.txt beside the library, or the file the MESHOPT_HIP_BATCH_SHAPE environment
 * variable names), unless a shape was set here first.  set: checks the shape (widths, the
 * wavefronts per SIMD each width's registers allow, workgroups per CU not growing over the
 * classes, the last class one workgroup per CU) and applies it to later calls; three NULL arrays
 * restore the default; returns 0, or -1 with the reason in meshopt_hip_last_error.  get: the
 * shape in use and the vertex and index caps it gives each class on the current device (any
 * array may be NULL); returns 0, or -1 on a HIP error.  A mesh's result never depends on the shape. */
int meshopt_hip_batch_set_shape(const unsigned int* per_cu, const unsigned int* threads, const unsigned int* cursor);
int meshopt_hip_batch_get_shape(unsigned int* per_cu, unsigned int* threads, unsigned int* cursor, unsigned int* vertex_caps, unsigned int* index_caps);

/* Serving.  A server runs simplify requests of any size as they arrive, on `slots` concurrent
 * device contexts (each its own HIP stream, arena and staging buffer), admitting them in arrival
 * order against a device-memory budget: a request starts once a slot is idle and the slot's
 * memory, next to what the other slots hold, fits the budget.  A small request without Sparse
 * or Prune (at most 1984 vertices and 5952 indices) takes the small requests queued behind it with
 * the same target error, options, lock presence, attribute count and weights into one batched call
 * on its slot.  Each result is exactly what meshopt_simplify /
 * meshopt_simplifyWithAttributes (src/meshoptimizer.h:513, 537) returns for the request alone.
 * This interface has no counterpart in the reference. */
typedef struct meshopt_hip_server meshopt_hip_server;

typedef struct meshopt_hip_request
{
	unsigned int* destination; /* host array with room for index_count indices */
	const unsigned int* indices;
	size_t index_count;
	const float* vertex_positions;
	size_t vertex_count;
	size_t vertex_positions_stride;
	const float* vertex_attributes; /* NULL, or as in meshopt_simplifyWithAttributes */
	size_t vertex_attributes_stride;
	const float* attribute_weights; /* copied at submit */
	size_t attribute_count;
	const unsigned char* vertex_lock; /* may be NULL */
	size_t target_index_count;
	float target_error;
	unsigned int options;
} meshopt_hip_request;

typedef struct meshopt_hip_response
{
	size_t index_count; /* indices written to the request's destination; (size_t)-1 on failure */
	float result_error;
	int slot;
	double submit_ms, start_ms, finish_ms; /* server clock (meshopt_hip_server_clock_ms) */
	size_t device_bytes;                   /* the admission estimate of the request */
	int batched; /* requests simplified by the same batched call (1: the plain path, alone) */
} meshopt_hip_response;

/* Returns NULL on failure (meshopt_hip_last_error). */
meshopt_hip_server* meshopt_hip_server_create(size_t memory_budget_bytes, int slots);
/* Queues a request; its arrays must stay valid until the ticket is waited on.  Returns the
 * ticket, or 0 if the request is invalid or needs more than the whole budget. */
size_t meshopt_hip_server_submit(meshopt_hip_server* server, const meshopt_hip_request* request);
/* Blocks until the ticket's request has finished; fills *response and frees the ticket.  Returns 0,
 * or -1 on failure (also for an unknown ticket, or one another thread already waits on). */
int meshopt_hip_server_wait(meshopt_hip_server* server, size_t ticket, meshopt_hip_response* response);
double meshopt_hip_server_clock_ms(const meshopt_hip_server* server);
/* Finishes every queued request, then frees the server's contexts. */
void meshopt_hip_server_destroy(meshopt_hip_server* server);

/* Text of the last HIP error seen by the calling thread (empty string if none). */
const char* meshopt_hip_last_error(void);

/* Releases the device arenas and streams held by the library (they are re-created on demand). */
void meshopt_hip_release(void);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* MESHOPT_HIP_H */
