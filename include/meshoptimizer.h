/**
 * meshoptimizer.h -- drop-in declarations for the simplifier entry points served by
 * libmeshopt_hip.so (gfx950 / MI355X).
 *
 * Same C ABI, enum values and argument meaning as meshoptimizer 1.2 (MESHOPTIMIZER_VERSION
 * 1020), simplification section.  A caller that links libmeshopt_hip.so instead of
 * libmeshoptimizer gets the same output index buffer, index count and result_error, bit for
 * bit.  Each declaration names the reference interface it replaces.
 *
 * The header is C89-clean (the reference checks that with demo/ansi.c:1).
 */
#ifndef MESHOPT_HIP_MESHOPTIMIZER_H
#define MESHOPT_HIP_MESHOPTIMIZER_H

#include <stddef.h>

/* reference: src/meshoptimizer.h:15 */
#define MESHOPTIMIZER_VERSION 1020

#ifndef MESHOPTIMIZER_API
#define MESHOPTIMIZER_API
#endif

#ifndef MESHOPTIMIZER_ALLOC_CALLCONV
#define MESHOPTIMIZER_ALLOC_CALLCONV
#endif

#ifdef __cplusplus
extern "C"
{
#endif

/* Simplification options; values of src/meshoptimizer.h:466-482 */
enum
{
	meshopt_SimplifyLockBorder = 1 << 0,      /* keep vertices on the topological border fixed */
	meshopt_SimplifySparse = 1 << 1,          /* indices reference a sparse subset of the vertex buffer */
	meshopt_SimplifyErrorAbsolute = 1 << 2,   /* target_error / result_error are absolute, not relative to the extent */
	meshopt_SimplifyPrune = 1 << 3,           /* remove small disconnected components while simplifying */
	meshopt_SimplifyRegularize = 1 << 4,      /* stronger regularisation of triangle shapes */
	meshopt_SimplifyPermissive = 1 << 5,      /* allow collapses across attribute discontinuities (not Protect-ed ones) */
	meshopt_SimplifyRegularizeLight = 1 << 6  /* lighter regularisation of triangle shapes */
};

/* Per-vertex flags for vertex_lock; values of src/meshoptimizer.h:487-495 */
enum
{
	meshopt_SimplifyVertex_Lock = 1 << 0,     /* the vertex never moves */
	meshopt_SimplifyVertex_Protect = 1 << 1,  /* keep the attribute discontinuity here (with Permissive) */
	meshopt_SimplifyVertex_Priority = 1 << 2  /* stronger regularisation: the vertex is more likely to survive */
};

/**
 * Quadric edge-collapse simplifier.
 * Replaces meshopt_simplify, src/meshoptimizer.h:513 (implementation src/simplifier.cpp:2628-2633).
 * destination needs room for index_count indices and may alias indices; vertex_positions holds
 * float3 positions in the first 12 bytes of each vertex (stride 12..256, a multiple of 4);
 * result_error may be NULL.  Returns the number of indices written.
 */
MESHOPTIMIZER_API size_t meshopt_simplify(unsigned int* destination, const unsigned int* indices, size_t index_count, const float* vertex_positions, size_t vertex_count, size_t vertex_positions_stride, size_t target_index_count, float target_error, unsigned int options, float* result_error);

/**
 * Quadric edge-collapse simplifier with an attribute metric and per-vertex locks.
 * Replaces meshopt_simplifyWithAttributes, src/meshoptimizer.h:537 (implementation
 * src/simplifier.cpp:2635-2640).  attribute_count <= 32; attributes with weight 0 are ignored;
 * vertex_lock may be NULL.
 */
MESHOPTIMIZER_API size_t meshopt_simplifyWithAttributes(unsigned int* destination, const unsigned int* indices, size_t index_count, const float* vertex_positions, size_t vertex_count, size_t vertex_positions_stride, const float* vertex_attributes, size_t vertex_attributes_stride, const float* attribute_weights, size_t attribute_count, const unsigned char* vertex_lock, size_t target_index_count, float target_error, unsigned int options, float* result_error);

/**
 * Edge-collapse simplifier that also moves the surviving vertices: after the collapses, each
 * remaining vertex that is not locked, on a border or on an attribute seam is solved to the
 * optimum of its accumulated quadric (with attributes: of every wedge's attribute quadric, under
 * a volume-preservation constraint), and the attributes are re-evaluated there.  The index
 * buffer is simplified in place; vertex_positions and vertex_attributes are updated in place
 * (externally locked vertices are left untouched).  Replaces meshopt_simplifyWithUpdate,
 * src/meshoptimizer.h:560 (implementation src/simplifier.cpp:2642-2645, 2575-2599).
 */
MESHOPTIMIZER_API size_t meshopt_simplifyWithUpdate(unsigned int* indices, size_t index_count, float* vertex_positions, size_t vertex_count, size_t vertex_positions_stride, float* vertex_attributes, size_t vertex_attributes_stride, const float* attribute_weights, size_t attribute_count, const unsigned char* vertex_lock, size_t target_index_count, float target_error, unsigned int options, float* result_error);

/**
 * Sloppy simplifier: vertices snap to a regular grid whose size is searched for the target, each
 * grid cell keeps its lowest-error vertex, collapsed and duplicate triangles are dropped.
 * Replaces meshopt_simplifySloppy, src/meshoptimizer.h:576 (implementation
 * src/simplifier.cpp:2647-2778).  destination needs room for index_count indices and may alias
 * indices; vertex_lock may be NULL; result_error may be NULL.  Returns the number of indices written.
 */
MESHOPTIMIZER_API size_t meshopt_simplifySloppy(unsigned int* destination, const unsigned int* indices, size_t index_count, const float* vertex_positions, size_t vertex_count, size_t vertex_positions_stride, const unsigned char* vertex_lock, size_t target_index_count, float target_error, float* result_error);

/**
 * Largest side of the positions' bounding box: the factor between absolute and relative error.
 * Replaces meshopt_simplifyScale, src/meshoptimizer.h:611 (implementation src/simplifier.cpp:2938-2948).
 */
MESHOPTIMIZER_API float meshopt_simplifyScale(const float* vertex_positions, size_t vertex_count, size_t vertex_positions_stride);

/**
 * Allocation callbacks for host memory: the index adapters of the C++ overloads below (16-bit index
 * buffers widened for the call) take their buffers through them, and so would any host temporary of
 * the library (its entry points stage inputs in device memory, which the library manages itself,
 * and make no host allocation per call).  Every block is freed before the call that took it returns.
 * Replaces meshopt_setAllocator, src/meshoptimizer.h:1042 (implementation src/allocator.cpp:14-21).
 */
MESHOPTIMIZER_API void meshopt_setAllocator(void* (MESHOPTIMIZER_ALLOC_CALLCONV* allocate)(size_t), void (MESHOPTIMIZER_ALLOC_CALLCONV* deallocate)(void*));

#ifdef __cplusplus
} /* extern "C" */
#endif

#ifdef __cplusplus
/* The callbacks meshopt_setAllocator sets (operator new / delete until then), shared by the
 * library and its callers: the library exports storage(), as the reference does when built with
 * MESHOPTIMIZER_ALLOC_EXPORT (src/meshoptimizer.h:1194-1244, src/allocator.cpp:6-12). */
class meshopt_Allocator
{
public:
	struct Storage
	{
		void* (MESHOPTIMIZER_ALLOC_CALLCONV* allocate)(size_t);
		void (MESHOPTIMIZER_ALLOC_CALLCONV* deallocate)(void*);
	};

	MESHOPTIMIZER_API static Storage& storage();
};
#endif

#if defined(__cplusplus) && !defined(MESHOPTIMIZER_NO_WRAPPERS)
/* Index-type adapters: same role as the templated overloads of src/meshoptimizer.h:1514-1538.
 * A 16-bit index buffer is widened into a block from meshopt_Allocator (the role of
 * meshopt_IndexAdapter, src/meshoptimizer.h:1249-1296), simplified through the C entry point, and
 * narrowed back into `result` when the adapter goes out of scope; 32-bit buffers are used as they
 * are. */
template <typename T, bool Wide = sizeof(T) == sizeof(unsigned int)>
struct meshopt_IndexAdapter
{
	T* result;
	unsigned int* data;
	size_t count;

	meshopt_IndexAdapter(T* result_, const T* input, size_t count_)
	    : result(result_)
	    , data(NULL)
	    , count(count_)
	{
		const size_t bytes = count > size_t(-1) / sizeof(unsigned int) ? size_t(-1) : count * sizeof(unsigned int);
		data = static_cast<unsigned int*>(meshopt_Allocator::storage().allocate(bytes));
		for (size_t i = 0; input && i < count; ++i)
			data[i] = input[i];
	}

	~meshopt_IndexAdapter()
	{
		for (size_t i = 0; result && i < count; ++i)
			result[i] = T(data[i]);
		meshopt_Allocator::storage().deallocate(data);
	}

private:
	meshopt_IndexAdapter(const meshopt_IndexAdapter&);
	meshopt_IndexAdapter& operator=(const meshopt_IndexAdapter&);
};

template <typename T>
struct meshopt_IndexAdapter<T, true>
{
	unsigned int* data;

	meshopt_IndexAdapter(T* result, const T* input, size_t)
	    : data(reinterpret_cast<unsigned int*>(result ? result : const_cast<T*>(input)))
	{
	}
};

template <typename T>
inline size_t meshopt_simplify(T* destination, const T* indices, size_t index_count, const float* vertex_positions, size_t vertex_count, size_t vertex_positions_stride, size_t target_index_count, float target_error, unsigned int options = 0, float* result_error = NULL)
{
	meshopt_IndexAdapter<T> in(NULL, indices, index_count);
	meshopt_IndexAdapter<T> out(destination, NULL, index_count);
	return meshopt_simplify(out.data, in.data, index_count, vertex_positions, vertex_count, vertex_positions_stride, target_index_count, target_error, options, result_error);
}

template <typename T>
inline size_t meshopt_simplifyWithAttributes(T* destination, const T* indices, size_t index_count, const float* vertex_positions, size_t vertex_count, size_t vertex_positions_stride, const float* vertex_attributes, size_t vertex_attributes_stride, const float* attribute_weights, size_t attribute_count, const unsigned char* vertex_lock, size_t target_index_count, float target_error, unsigned int options = 0, float* result_error = NULL)
{
	meshopt_IndexAdapter<T> in(NULL, indices, index_count);
	meshopt_IndexAdapter<T> out(destination, NULL, index_count);
	return meshopt_simplifyWithAttributes(out.data, in.data, index_count, vertex_positions, vertex_count, vertex_positions_stride, vertex_attributes, vertex_attributes_stride, attribute_weights, attribute_count, vertex_lock, target_index_count, target_error, options, result_error);
}

template <typename T>
inline size_t meshopt_simplifyWithUpdate(T* indices, size_t index_count, float* vertex_positions, size_t vertex_count, size_t vertex_positions_stride, float* vertex_attributes, size_t vertex_attributes_stride, const float* attribute_weights, size_t attribute_count, const unsigned char* vertex_lock, size_t target_index_count, float target_error, unsigned int options = 0, float* result_error = NULL)
{
	meshopt_IndexAdapter<T> inout(indices, indices, index_count);
	return meshopt_simplifyWithUpdate(inout.data, index_count, vertex_positions, vertex_count, vertex_positions_stride, vertex_attributes, vertex_attributes_stride, attribute_weights, attribute_count, vertex_lock, target_index_count, target_error, options, result_error);
}

template <typename T>
inline size_t meshopt_simplifySloppy(T* destination, const T* indices, size_t index_count, const float* vertex_positions, size_t vertex_count, size_t vertex_positions_stride, const unsigned char* vertex_lock, size_t target_index_count, float target_error, float* result_error = NULL)
{
	meshopt_IndexAdapter<T> in(NULL, indices, index_count);
	meshopt_IndexAdapter<T> out(destination, NULL, index_count);
	return meshopt_simplifySloppy(out.data, in.data, index_count, vertex_positions, vertex_count, vertex_positions_stride, vertex_lock, target_index_count, target_error, result_error);
}

template <typename T>
inline size_t meshopt_simplifySloppy(T* destination, const T* indices, size_t index_count, const float* vertex_positions, size_t vertex_count, size_t vertex_positions_stride, size_t target_index_count, float target_error, float* result_error = NULL)
{
	return meshopt_simplifySloppy(destination, indices, index_count, vertex_positions, vertex_count, vertex_positions_stride, (const unsigned char*)NULL, target_index_count, target_error, result_error);
}
#endif

#endif /* MESHOPT_HIP_MESHOPTIMIZER_H */
