// bucket_order.h -- stable counting sort of a device list by a small integer key.
//
// Used to form waves of similar work: entries whose previous step took a
// similar number of passes (Richards sub-steps, canopy iterations) end up next
// to each other.  Three kernels, no host work, so it can be captured in a
// graph: per block the count of each key, one exclusive scan over
// [key][block], then a scatter that keeps list order within a key (a block's
// entries are placed wave by wave, lanes in order).  The number of entries
// may be known only on the device (count != nullptr: *count entries, at most
// n_max).  Key: a functor with `int key(int i)` in [0, BO_KEYS) and
// `int32_t value(int i)`, the entry to place, for list position i.
#pragma once

#include "common.h"

namespace ctsm {

constexpr int BO_KEYS = 64;
constexpr int BO_THREADS = 256;
constexpr int BO_PASSES = 8;                      // list entries per thread
constexpr int BO_SPAN = BO_THREADS * BO_PASSES;   // list entries per block

inline int bucket_order_blocks(int n_max) { return (n_max + BO_SPAN - 1) / BO_SPAN; }
// Workspace of the sort: the [key][block] table (the output list is the caller's).
inline size_t bucket_order_table_bytes(int n_max)
{
    return (size_t)BO_KEYS * (size_t)bucket_order_blocks(n_max) * sizeof(int32_t);
}

template <class Key>
__global__ void __launch_bounds__(BO_THREADS) k_bo_count(Key K, int n_max, const int32_t *count, int32_t *table)
{
    __shared__ int cnt[BO_KEYS];
    if (threadIdx.x < BO_KEYS) cnt[threadIdx.x] = 0;
    __syncthreads();
    const int n = count ? min(*count, n_max) : n_max;
    for (int p = 0; p < BO_PASSES; ++p) {
        const int i = blockIdx.x * BO_SPAN + p * BO_THREADS + threadIdx.x;
        if (i < n) atomicAdd(&cnt[K.key(i)], 1);
    }
    __syncthreads();
    if (threadIdx.x < BO_KEYS) table[(long)threadIdx.x * gridDim.x + blockIdx.x] = cnt[threadIdx.x];
}

// exclusive prefix sum of table[0 .. m) in one block
static __global__ void __launch_bounds__(1024) k_bo_scan(int32_t *table, int m)
{
    __shared__ int part[1024];
    const int t = threadIdx.x;
    const int per = (m + 1023) / 1024;
    const int lo = t * per, hi = min(lo + per, m);
    int s = 0;
    for (int k = lo; k < hi; ++k) s += table[k];
    part[t] = s;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {
        const int v = t >= d ? part[t - d] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int run = part[t] - s;
    for (int k = lo; k < hi; ++k) {
        const int v = table[k];
        table[k] = run;
        run += v;
    }
}

template <class Key>
__global__ void __launch_bounds__(BO_THREADS) k_bo_scatter(Key K, int n_max, const int32_t *count,
                                                          const int32_t *__restrict__ table, int32_t *out)
{
    __shared__ int next[BO_KEYS];
    if (threadIdx.x < BO_KEYS) next[threadIdx.x] = table[(long)threadIdx.x * gridDim.x + blockIdx.x];
    const int n = count ? min(*count, n_max) : n_max;
    const int lane = threadIdx.x & 63;
    const unsigned long long below = (1ull << lane) - 1ull;
    for (int p = 0; p < BO_PASSES; ++p) {
        const int i = blockIdx.x * BO_SPAN + p * BO_THREADS + threadIdx.x;
        const int key = i < n ? K.key(i) : -1;
        const int32_t v = i < n ? K.value(i) : 0;
        for (int w = 0; w < BO_THREADS / 64; ++w) {
            __syncthreads();
            if ((int)(threadIdx.x >> 6) != w) continue;
            // the wave's keys one at a time; lanes in order = list order
            unsigned long long todo = __ballot(key >= 0);
            while (todo) {
                const int first = __ffsll((unsigned long long)todo) - 1;
                const int k = __shfl(key, first);
                const unsigned long long same = __ballot(key == k);
                if (key == k) out[next[k] + __popcll(same & below)] = v;
                if (lane == first) next[k] += __popcll(same);
                todo &= ~same;
            }
        }
    }
}

// Enqueues the sort of n_max list positions (or *count of them) into out,
// with `table` of bucket_order_table_bytes(n_max) bytes.
template <class Key>
void launch_bucket_order(hipStream_t st, const Key &K, int n_max, const int32_t *count, int32_t *table,
                         int32_t *out)
{
    if (n_max <= 0) return;
    const int nblk = bucket_order_blocks(n_max);
    hipLaunchKernelGGL(k_bo_count<Key>, dim3(nblk), dim3(BO_THREADS), 0, st, K, n_max, count, table);
    hipLaunchKernelGGL(k_bo_scan, dim3(1), dim3(1024), 0, st, table, BO_KEYS * nblk);
    hipLaunchKernelGGL(k_bo_scatter<Key>, dim3(nblk), dim3(BO_THREADS), 0, st, K, n_max, count, table, out);
}

}  // namespace ctsm
