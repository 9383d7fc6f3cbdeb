/* raft_noalloc.h -- the allocation-free entry point of libraft.
 *
 * No counterpart in the reference (its GPU entries allocate and free on every call,
 * raft/raft_radon_gpu.cu:16-35).  It is for callers that cannot allocate device memory:
 * code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.
 */
#ifndef PYRAFT_AMD_RAFT_NOALLOC_H
#define PYRAFT_AMD_RAFT_NOALLOC_H

#include "raft_slantstack.h"

#ifdef __cplusplus
extern "C" {
#endif
#pragma GCC visibility push(default)

/* Bytes of the library's static device scratch array: two gather layouts of one
 * 2048 x 2048 slice (2 * 2048 * 2064 * 8 bytes) plus the row-slab partial sums of a
 * divided launch.  Every GPU the library is loaded on holds this array until the
 * process ends. */
#define RAFT_SS_NOALLOC_SCRATCH_BYTES (72u * 1024u * 1024u)

/* raft_ss_radon_range without allocating: the same kernels, with the workspace (the
 * image's two gather layouts and, for launches of few angles, the slabs' partial ray
 * sums) in the static device array instead of stream-ordered allocations.  The caller
 * passes the plan, input and output; the result is bit-identical to raft_ss_radon_range.
 * Calls no hipMalloc, hipFree or other allocating routine; the device address of the
 * array is looked up once per device on the first call there.
 * Workspace need: 16 * nslices * n * pitch bytes (pitch = n + 2 rounded up to 16) plus
 * the slab sums; a call needing more than RAFT_SS_NOALLOC_SCRATCH_BYTES returns
 * RAFT_SS_EINVAL with a message and launches nothing.
 * All calls on one GPU share that GPU's array, which keeps the last call's contents:
 * calls that overlap on one GPU (two streams, or two host threads) race and must be
 * serialised by the caller.  Nothing in the array needs to start at zero. */
int raft_ss_radon_noalloc(const raft_ss_plan *plan, const float *d_image, float *d_out, int angle_major, int j0,
                          int j1, int nslices, void *stream);

#pragma GCC visibility pop
#ifdef __cplusplus
}
#endif

#endif
