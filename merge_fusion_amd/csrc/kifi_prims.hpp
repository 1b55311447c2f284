// kifi_prims.hpp -- wave64 / workgroup scan and reduce primitives (device).
//
// Written for CDNA4's 64-lane wavefronts: a workgroup of 256 threads is four
// waves; cross-wave steps go through a 4-entry LDS array.  The reference's
// unbuilt CUDA primitives (cuda/warp.cuh, cuda/block.cuh) hard-code 32-lane
// warps and 32-bit masks; these are independent wave64 designs.
#pragma once

#include "kifi_internal.hpp"

namespace kifi {

template <typename T>
__device__ __forceinline__ T wave_inclusive_scan(T v)
{
    const int lane = __lane_id();
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        T t = __shfl_up(v, o, kWave);
        if (lane >= o)
            v += t;
    }
    return v;
}

template <typename T>
__device__ __forceinline__ T wave_reduce_sum(T v)
{
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1)
        v += __shfl_xor(v, o, kWave);
    return v;
}

// Exclusive scan over a workgroup of NT threads (NT multiple of 64, <= 1024).
// `tmp` must hold NT/64 + 1 entries; returns the exclusive prefix and the total.
template <int NT, typename T>
__device__ __forceinline__ T block_exclusive_scan(T v, T *tmp, T &total)
{
    constexpr int NW = NT / kWave;
    const int w = threadIdx.x / kWave, lane = __lane_id();
    T inc = wave_inclusive_scan(v);
    if (lane == kWave - 1)
        tmp[w] = inc;
    __syncthreads();
    if (threadIdx.x == 0) {
        T acc = 0;
        for (int i = 0; i < NW; ++i) {
            T t = tmp[i];
            tmp[i] = acc;
            acc += t;
        }
        tmp[NW] = acc;
    }
    __syncthreads();
    T r = tmp[w] + inc - v;
    total = tmp[NW];
    __syncthreads();
    return r;
}

template <int NT, typename T>
__device__ __forceinline__ T block_reduce_sum(T v, T *tmp)
{
    constexpr int NW = NT / kWave;
    const int w = threadIdx.x / kWave, lane = __lane_id();
    v = wave_reduce_sum(v);
    if (lane == 0)
        tmp[w] = v;
    __syncthreads();
    T acc = 0;
    for (int i = 0; i < NW; ++i)
        acc += tmp[i];
    __syncthreads();
    return acc;
}

template <int NT, typename T>
__device__ __forceinline__ T block_reduce_max(T v, T *tmp)
{
    constexpr int NW = NT / kWave;
    const int w = threadIdx.x / kWave, lane = __lane_id();
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) {
        const T u = __shfl_xor(v, o, kWave);
        v = u > v ? u : v;
    }
    if (lane == 0)
        tmp[w] = v;
    __syncthreads();
    T acc = tmp[0];
    for (int i = 1; i < NW; ++i)
        acc = tmp[i] > acc ? tmp[i] : acc;
    __syncthreads();
    return acc;
}

// Number of elements of sorted s[0, n) that are < x (lower bound).
template <typename K>
__device__ __forceinline__ uint32_t lds_lower_bound(const K *s, uint32_t n, K x)
{
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (s[mid] < x)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// Single-workgroup exclusive scan of a device array of n entries in place
// (n up to a few hundred thousand); returns the total through *total.
template <int NT, typename T, typename TT>
__device__ void block_scan_array(T *a, size_t n, TT *total)
{
    __shared__ TT tmp[NT / kWave + 1];
    TT carry = 0;
    constexpr int ITEMS = 8;
    for (size_t start = 0; start < n; start += (size_t)NT * ITEMS) {
        TT loc[ITEMS];
        TT sum = 0;
        const size_t mybase = start + (size_t)threadIdx.x * ITEMS;
#pragma unroll
        for (int q = 0; q < ITEMS; ++q) {
            loc[q] = (mybase + q < n) ? (TT)a[mybase + q] : (TT)0;
            sum += loc[q];
        }
        TT tot;
        TT ex = block_exclusive_scan<NT>(sum, tmp, tot);
        TT run = carry + ex;
#pragma unroll
        for (int q = 0; q < ITEMS; ++q) {
            if (mybase + q < n)
                a[mybase + q] = (T)run;
            run += loc[q];
        }
        carry += tot;
    }
    if (threadIdx.x == 0 && total)
        *total = carry;
}

} // namespace kifi
