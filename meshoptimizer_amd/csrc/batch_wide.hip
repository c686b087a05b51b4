// The batched kernel's device code built for 1024-thread workgroups (16 wavefronts), for a size
// class whose workgroup takes a whole CU's LDS: with 256 threads each SIMD would hold a single
// wavefront.  The device code is the same source as batch.hip's, in its own inline namespace.
#define MHIP_BATCH_THREADS 1024
#include "batch_kernels.h"
