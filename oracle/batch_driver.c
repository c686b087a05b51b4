/*
 * Batch driver for the CPU baselines (test and benchmark infrastructure only): runs any function
 * with meshopt_simplify's signature (src/meshoptimizer.h:513) -- the reference library built in
 * oracle/_ref or the C restatement -- over a packed batch of independent meshes, one mesh per
 * work item, on a pool of POSIX threads with dynamic scheduling.  Mesh p owns indices
 * [index_offsets[p], index_offsets[p+1]) and vertices [vertex_offsets[p], vertex_offsets[p+1]);
 * its result is written at destination + index_offsets[p].
 */
#include <pthread.h>
#include <stdatomic.h>
#include <stddef.h>
#include <stdlib.h>

typedef size_t (*simplify_fn)(unsigned int*, const unsigned int*, size_t, const float*, size_t, size_t, size_t, float, unsigned int, float*);

typedef struct
{
	simplify_fn fn;
	unsigned int* destination;
	unsigned int* result_counts;
	float* result_errors;
	const unsigned int* indices;
	const unsigned int* index_offsets;
	const float* positions;
	const unsigned int* vertex_offsets;
	size_t stride;
	const unsigned int* targets;
	size_t count;
	float target_error;
	unsigned int options;
	atomic_size_t next;
} batch_job;

static void* batch_worker(void* arg)
{
	batch_job* j = (batch_job*)arg;
	for (;;)
	{
		size_t p = atomic_fetch_add(&j->next, 1);
		if (p >= j->count)
			break;
		size_t ib = j->index_offsets[p], ni = j->index_offsets[p + 1] - ib;
		size_t vb = j->vertex_offsets[p], nv = j->vertex_offsets[p + 1] - vb;
		float e = 0;
		size_t r = j->fn(j->destination + ib, j->indices + ib, ni, j->positions + vb * (j->stride / sizeof(float)), nv, j->stride,
		                 j->targets[p], j->target_error, j->options, &e);
		j->result_counts[p] = (unsigned int)r;
		if (j->result_errors)
			j->result_errors[p] = e;
	}
	return NULL;
}

/* Returns 0, or -1 if a thread could not be started. */
int batch_simplify(void* fn, int threads, unsigned int* destination, unsigned int* result_counts, float* result_errors, const unsigned int* indices,
                   const unsigned int* index_offsets, const float* positions, const unsigned int* vertex_offsets, size_t stride,
                   const unsigned int* targets, size_t count, float target_error, unsigned int options)
{
	batch_job j;
	j.fn = (simplify_fn)fn;
	j.destination = destination;
	j.result_counts = result_counts;
	j.result_errors = result_errors;
	j.indices = indices;
	j.index_offsets = index_offsets;
	j.positions = positions;
	j.vertex_offsets = vertex_offsets;
	j.stride = stride;
	j.targets = targets;
	j.count = count;
	j.target_error = target_error;
	j.options = options;
	atomic_init(&j.next, 0);

	if (threads <= 1)
	{
		batch_worker(&j);
		return 0;
	}

	pthread_t* tid = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)threads);
	if (!tid)
		return -1;
	int started = 0, rc = 0;
	for (; started < threads; ++started)
		if (pthread_create(&tid[started], NULL, batch_worker, &j) != 0)
		{
			rc = -1;
			break;
		}
	for (int t = 0; t < started; ++t)
		pthread_join(tid[t], NULL);
	free(tid);
	return rc;
}
