/* Host-only stand-in for the few HIP runtime calls the checkpoint writer
 * makes, so that its file format can be tested on a machine without a GPU:
 * pinned allocations are malloc, copies are memcpy (synchronous), events and
 * streams do nothing.  Test infrastructure only. */
#pragma once
#include <cstdlib>
#include <cstring>
typedef int hipError_t;
typedef void *hipStream_t;
typedef void *hipEvent_t;
#define hipSuccess 0
#define hipMemcpyDefault 4
#define hipHostMallocDefault 0
#define hipEventDisableTiming 2
inline hipError_t hipHostMalloc(void **p, size_t n, int) { *p = malloc(n); return *p ? 0 : 1; }
inline hipError_t hipHostFree(void *p) { free(p); return 0; }
inline hipError_t hipEventCreateWithFlags(hipEvent_t *e, int) { *e = (void *)1; return 0; }
inline hipError_t hipEventDestroy(hipEvent_t) { return 0; }
inline hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return 0; }
inline hipError_t hipEventSynchronize(hipEvent_t) { return 0; }
inline hipError_t hipStreamSynchronize(hipStream_t) { return 0; }
inline hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, int, hipStream_t) { memcpy(d, s, n); return 0; }
