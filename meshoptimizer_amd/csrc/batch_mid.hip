// The batched kernel's device code built for 512-thread workgroups (8 wavefronts), a width the
// batch shape may give a size class of two or three workgroups per CU.  The device code is the
// same source as batch.hip's, in its own inline namespace.
#define MHIP_BATCH_THREADS 512
#include "batch_kernels.h"
