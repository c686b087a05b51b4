// The batched kernel's device code built for 128-thread workgroups (2 wavefronts): twelve of them
// share a CU within the 80-VGPR budget (six wavefronts per SIMD), so a size class of small meshes
// (about 13 KB of LDS each) runs twice as many meshes side by side as at 256 threads, and only one
// wavefront waits while wavefront 0 runs the greedy commit.  The device code is the same source
// as batch.hip's, in its own inline namespace.
#define MHIP_BATCH_THREADS 128
#include "batch_kernels.h"
